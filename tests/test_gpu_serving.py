"""Continuous-batching server on the GPU: every job's accumulator is the
same, bit for bit, whether the job renders alone (PathTracer.refine) or
inside a mix of scenes, resolutions, spp and bounce limits that share
wavefront launches; admission by memory budget queues jobs and admits
them as others finish."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _opts(w, h, spp, bounces, ambient):
    import oxr
    o = oxr.RenderOptions(width=w, height=h, ambient=ambient)
    o.path_trace.samples_per_pixel = spp
    o.path_trace.max_bounces = bounces
    return o


def _alone(scene, opts):
    import oxr
    pt = oxr.PathTracer(0)
    pt.sync(scene, opts)
    pt.refine(opts.path_trace.samples_per_pixel)
    acc, sq = pt.accum(with_sum_sq=True)
    pt.close()
    return acc, sq


@pytest.fixture(scope="module")
def scenes():
    from oxr import scenes as S
    return [S.cornell(), S.material_sweep(segments=16, rings=8, floor_half=6), S.triangle_soup(11, 3000, n_emit=3)]


MIX = [  # (scene, w, h, spp, bounces)
    (0, 64, 64, 16, 8),
    (1, 96, 80, 7, 4),
    (2, 50, 33, 9, 6),
    (0, 17, 128, 4, 5),
    (1, 64, 64, 33, 8),
    (2, 128, 128, 5, 8),
    (0, 1, 1, 3, 8),
]


def _run_mix(server, ids, scenes, mix):
    jobs = []
    for si, w, h, spp, b in mix:
        s = scenes[si]
        jid = server.submit(ids[si], s.camera.resolve(w, h), _opts(w, h, spp, b, s.ambient))
        jobs.append(jid)
    max_active = 0
    for _ in range(10000):
        a, _ = server.step()
        max_active = max(max_active, a)
        if all(server.status(j).state == 2 for j in jobs):
            break
    return jobs, max_active


def test_mix_equals_alone(scenes, monkeypatch):
    from oxr.serve import Server
    # A small round budget (paths per wavefront batch) splits every job
    # over several rounds, so jobs enter and leave while others render.
    monkeypatch.setenv("OXR_PATH_BUDGET", "65536")
    srv = Server(0)
    ids = [srv.add_scene(s) for s in scenes]
    jobs, max_active = _run_mix(srv, ids, scenes, MIX)
    assert max_active >= 3  # jobs overlapped inside shared launches
    assert srv.stats()["rounds"] >= 4
    for jid, (si, w, h, spp, b) in zip(jobs, MIX):
        st = srv.status(jid)
        assert st.state == 2 and st.samples == spp
        acc, sq = srv.accum(jid, with_sum_sq=True)
        ref, rsq = _alone(scenes[si], _opts(w, h, spp, b, scenes[si].ambient))
        assert np.array_equal(acc, ref), (si, w, h, spp, b)
        assert np.array_equal(sq, rsq), (si, w, h, spp, b)
        srv.release(jid)
    srv.close()


def test_small_budget_queues_and_admits(scenes):
    from oxr.serve import Server
    # The budget covers the resident scenes, the wavefront buffers and the
    # accumulators: room for the scenes and about two 64x64 accumulators
    # (7 doubles / pixel) before the first round allocates wavefronts.
    probe = Server(0)
    for s in scenes:
        probe.add_scene(s)
    base = probe.stats()["bytes_resident"]
    probe.close()
    srv = Server(0, mem_budget=base + 2 * 64 * 64 * 7 * 8 + 1024)
    ids = [srv.add_scene(s) for s in scenes]
    mix = [(0, 64, 64, 8, 8), (1, 64, 64, 8, 4), (2, 64, 64, 8, 6), (0, 64, 64, 6, 3)]
    jobs = []
    for si, w, h, spp, b in mix:
        jobs.append(srv.submit(ids[si], scenes[si].camera.resolve(w, h), _opts(w, h, spp, b, scenes[si].ambient)))
    srv.step()
    # Two accumulators fit; the other two jobs wait in the queue.
    assert [srv.status(j).state for j in jobs[2:]] == [0, 0]
    assert all(srv.status(j).state != 0 for j in jobs[:2])
    assert srv.stats()["queued"] == 2
    done = set()
    for _ in range(1000):
        srv.step()
        for j in jobs:
            if j not in done and srv.status(j).state == 2:
                done.add(j)
                srv.release(j)  # frees its accumulator, lets a queued job in
        if len(done) == len(jobs):
            break
    assert len(done) == len(jobs)
    for jid, (si, w, h, spp, b) in zip(jobs, mix):
        assert srv.status(jid).t_admit_ms >= srv.status(jid).t_submit_ms
    # Released jobs have no accumulator any more.
    import oxr
    with pytest.raises((oxr.BackendError, oxr.InvalidOptions)):
        srv.accum(jobs[0])
    srv.close()


def test_request_loop_reports_latency(scenes):
    from oxr.serve import Server, make_trace, run_requests
    srv = Server(0)
    ids = [srv.add_scene(s) for s in scenes]
    trace = make_trace(n=24, lam=200.0, n_scenes=3, seed=3)
    for r in trace:  # keep the test short
        r.width = r.height = min(r.width, 256)
        r.spp = min(r.spp, 16)
    out = run_requests(srv, ids, scenes, trace)
    assert out["completed"] == 24
    assert out["p50_ms"] <= out["p99_ms"]
    assert out["server"]["jobs_done"] == 24
    srv.close()


def test_prebuilt_fast_scene_job_equals_refine_and_is_freed(scenes):
    """A scene prepared off the owner thread with the linear (fast) BVH and
    uploaded with add_prebuilt renders the same bits as PathTracer.refine of
    the same scene with the SAH BVH (hits do not depend on the tree); once
    its job is released, remove_scene frees it and the id is refused."""
    import oxr
    from oxr import scenes as S
    from oxr.serve import Prebuilt, Server
    soup = S.triangle_soup(11, 30000, n_emit=3)
    srv = Server(0)
    pb = Prebuilt(soup, fast=True)
    sid = srv.add_prebuilt(pb)
    pb.close()
    assert srv.stats()["scenes"] == 1
    w, h, spp, b = 96, 80, 8, 6
    job = srv.submit(sid, soup.camera.resolve(w, h), _opts(w, h, spp, b, soup.ambient))
    for _ in range(100):
        srv.step()
        if srv.status(job).state == 2:
            break
    acc, sq = srv.accum(job, with_sum_sq=True)
    ref, rsq = _alone(soup, _opts(w, h, spp, b, soup.ambient))
    assert np.array_equal(acc, ref)
    assert np.array_equal(sq, rsq)
    srv.release(job)
    srv.remove_scene(sid)
    assert srv.stats()["scenes"] == 0
    with pytest.raises(oxr.InvalidOptions):
        srv.submit(sid, soup.camera.resolve(w, h), _opts(w, h, spp, b, soup.ambient))
    srv.close()


def test_rounds_after_a_removed_scene_keep_rendering(scenes):
    """A finished job's slot in the device job table keeps its (stale)
    entry until another job takes the slot. Here the job in slot 0 renders
    a large prebuilt soup, finishes, and its scene is freed while two jobs
    of another scene (slots 1 and 2) keep rendering in shared rounds; the
    shading threads past each queue's end must not read the freed scene
    through slot 0. Both jobs finish and equal refine alone."""
    from oxr import scenes as S
    from oxr.serve import Prebuilt, Server
    srv = Server(0)
    sid_a = srv.add_scene(scenes[0])
    soup = S.triangle_soup(21, 1_500_000, n_emit=2)
    pb = Prebuilt(soup, fast=True)
    sid_s = srv.add_prebuilt(pb)
    pb.close()
    j0 = srv.submit(sid_s, soup.camera.resolve(64, 64), _opts(64, 64, 2, 4, soup.ambient))
    srv.step()  # j0 takes slot 0
    mix = [(160, 96, 48, 6), (128, 128, 40, 8)]
    jobs = [srv.submit(sid_a, scenes[0].camera.resolve(w, h), _opts(w, h, spp, b, scenes[0].ambient))
            for w, h, spp, b in mix]
    for _ in range(200):
        srv.step()
        if srv.status(j0).state == 2:
            break
    assert srv.status(j0).state == 2
    srv.release(j0)
    srv.remove_scene(sid_s)
    assert srv.stats()["scenes"] == 1
    for _ in range(2000):
        srv.step()
        if all(srv.status(j).state == 2 for j in jobs):
            break
    for j, (w, h, spp, b) in zip(jobs, mix):
        assert srv.status(j).state == 2
        acc = srv.accum(j)
        ref, _ = _alone(scenes[0], _opts(w, h, spp, b, scenes[0].ambient))
        assert np.array_equal(acc, ref), (w, h, spp, b)
        srv.release(j)
    srv.close()
