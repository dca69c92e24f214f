"""Render serving over liboxr.so's continuous-batching server.

`Server` keeps scenes (BVH, light tables) and job accumulators resident on
one device. Jobs are admitted FIFO while their accumulator fits the memory
budget. `step()` runs one wavefront round over every active job; their
paths share the ray queues and the launches. `run_requests` replays a
request trace with Poisson arrivals against the wall clock and reports
p50/p99 latency beside throughput.

There is no counterpart in oxideav-render, which renders one scene per
`Renderer::render` call (src/lib.rs:136-167). Each job's result is
bit-identical to `PathTracer.refine(spp)` of the same scene and options.
"""
import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from ._lib import check, lib
from .pathtracer import RenderOptions, _pt_opts


QUEUED, ACTIVE, DONE = A.OXR_JOB_QUEUED, A.OXR_JOB_ACTIVE, A.OXR_JOB_DONE


class Server:
    def __init__(self, device=0, mem_budget=0):
        self._L = lib()
        h = C.c_void_p()
        check(self._L.oxr_srv_create(device, mem_budget, C.byref(h)), "oxr_srv_create")
        self._h = h
        self._descs = []
        self._dims = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.oxr_srv_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_scene(self, scene):
        d = scene.desc()
        sid = C.c_uint32(0)
        check(self._L.oxr_srv_add_scene(self._h, C.byref(d), C.byref(sid)), "add_scene")
        self._descs.append(d)
        return sid.value

    def add_prebuilt(self, pb):
        """Upload a Prebuilt scene (owner thread); returns its scene id."""
        sid = C.c_uint32(0)
        check(self._L.oxr_srv_add_prebuilt(self._h, pb.handle, C.byref(sid)), "add_prebuilt")
        return sid.value

    def remove_scene(self, scene_id):
        """Free a scene once no queued or active job uses it."""
        check(self._L.oxr_srv_remove_scene(self._h, scene_id), "remove_scene")

    def submit(self, scene_id, camera, opts: RenderOptions):
        jid = C.c_uint32(0)
        check(self._L.oxr_srv_submit(self._h, scene_id, C.byref(camera), C.byref(_pt_opts(opts)), C.byref(jid)),
              "submit")
        self._dims[jid.value] = (max(opts.width, 1), max(opts.height, 1))
        return jid.value

    def step(self):
        a = C.c_uint32(0)
        d = C.c_uint32(0)
        check(self._L.oxr_srv_step(self._h, C.byref(a), C.byref(d)), "step")
        return a.value, d.value

    def status(self, job):
        s = A.JobStatus()
        check(self._L.oxr_srv_job_status(self._h, job, C.byref(s)), "job_status")
        return s

    def accum(self, job, with_sum_sq=False):
        W, H = self._dims[job]
        acc = np.zeros((H, W, 4), np.float64)
        sq = np.zeros((H, W, 3), np.float64) if with_sum_sq else None
        dp = C.POINTER(C.c_double)
        check(self._L.oxr_srv_read_accum(self._h, job, acc.ctypes.data_as(dp),
                                         sq.ctypes.data_as(dp) if sq is not None else dp()), "read_accum")
        return (acc, sq) if with_sum_sq else acc

    def release(self, job):
        check(self._L.oxr_srv_release(self._h, job), "release")

    def stats(self):
        s = A.SrvStats()
        check(self._L.oxr_srv_stats_get(self._h, C.byref(s)), "stats")
        return {f: getattr(s, f) for f, _ in A.SrvStats._fields_}


class Prebuilt:
    """A scene prepared on the host (BVH, light tables, materials) by
    oxr_scene_prebuild: no device call, so any thread may build one while
    the server's owner thread keeps stepping. fast=True builds a linear
    (Morton-order) BVH with single-triangle leaves: about twice as fast to
    build as the binned SAH, and within 5% of its trace speed on random
    soups."""

    def __init__(self, scene, fast=True):
        self._L = lib()
        d = scene.desc()
        h = C.c_void_p()
        check(self._L.oxr_scene_prebuild(C.byref(d), A.OXR_BUILD_FAST if fast else 0, C.byref(h)), "scene_prebuild")
        self.handle = h
        self.num_tris = scene.num_tris

    def close(self):
        if getattr(self, "handle", None):
            self._L.oxr_prebuilt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


@dataclass
class Request:
    arrival_s: float
    scene: int          # index into the scene list given to run_requests (-1: a fresh soup)
    width: int
    height: int
    spp: int
    max_bounces: int
    soup_tris: int = 0  # a fresh seeded triangle soup of this many triangles (scene == -1)
    soup_seed: int = 0


def make_trace(n=512, lam=40.0, n_scenes=3, seed=0, soups=False):
    """Seeded request trace: Poisson arrivals at rate lam; scene uniform
    over n_scenes (with soups=True the last choice is a fresh seeded
    triangle soup per request, N log-uniform over 1k..2M triangles);
    resolution uniform over {128,256,512,1024}^2; spp uniform over
    {4,16,64,256}; max_bounces uniform over 4..8."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / lam, n)
    t = np.cumsum(gaps) - gaps[0]
    res = rng.choice([128, 256, 512, 1024], n)
    spp = rng.choice([4, 16, 64, 256], n)
    sc = rng.integers(0, n_scenes, n)
    b = rng.integers(4, 9, n)
    ntri = np.round(10.0 ** rng.uniform(3.0, np.log10(2e6), n)).astype(int)
    out = []
    for i in range(n):
        soup = soups and int(sc[i]) == n_scenes - 1
        out.append(Request(float(t[i]), -1 if soup else int(sc[i]), int(res[i]), int(res[i]), int(spp[i]), int(b[i]),
                           soup_tris=int(ntri[i]) if soup else 0, soup_seed=(seed * 100003 + i) & 0x7fffffff))
    return out


def _soup_job(r):
    """Worker-thread part of a soup request: generate the seeded soup and
    prepare it on the host (BVH, light tables)."""
    from .scenes import triangle_soup
    t = time.perf_counter()
    sc = triangle_soup(r.soup_seed, r.soup_tris, n_emit=1 + r.soup_seed % 8)
    pb = Prebuilt(sc, fast=True)
    return sc, pb, time.perf_counter() - t


def run_requests(server, scene_ids, scenes, trace, release=True, time_limit_s=None, build_workers=8):
    """Replay `trace` against the wall clock: submit each request when its
    arrival time comes, step the server continuously, record latency =
    completion - arrival. A fresh-soup request (scene -1) is generated and
    prepared on a pool of build threads while the server keeps stepping;
    its scene is uploaded and its job submitted when the build is done (the
    build counts in its latency), and the scene is freed with the job.
    Returns a summary dict."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    pending = list(range(len(trace)))
    jobs = {}
    done = {}
    building = {}  # future -> request index
    own_scene = {}  # job -> scene id to free with it
    build_s = []
    paths = 0
    pool = ThreadPoolExecutor(max_workers=build_workers) if any(r.scene < 0 for r in trace) else None

    def submit(i, sid, s):
        r = trace[i]
        cam = s.camera.resolve(r.width, r.height)
        opts = RenderOptions(width=r.width, height=r.height, ambient=getattr(s, "ambient", 0.0))
        opts.path_trace.samples_per_pixel = r.spp
        opts.path_trace.max_bounces = r.max_bounces
        j = server.submit(sid, cam, opts)
        jobs[j] = i
        return j

    try:
        while pending or building or len(done) < len(jobs):
            now = time.perf_counter() - t0
            if time_limit_s and now > time_limit_s:
                break
            while pending and trace[pending[0]].arrival_s <= now:
                i = pending.pop(0)
                r = trace[i]
                if r.scene < 0:
                    building[pool.submit(_soup_job, r)] = i
                else:
                    submit(i, scene_ids[r.scene], scenes[r.scene])
            for f in [f for f in building if f.done()]:
                i = building.pop(f)
                sc, pb, bs = f.result()
                build_s.append(bs)
                sid = server.add_prebuilt(pb)
                pb.close()
                own_scene[submit(i, sid, sc)] = sid
            active, _ = server.step()
            now = time.perf_counter() - t0
            for j, i in jobs.items():
                if j in done:
                    continue
                st = server.status(j)
                if st.state == DONE:
                    done[j] = now - trace[i].arrival_s
                    paths += trace[i].width * trace[i].height * trace[i].spp
                    if release:
                        server.release(j)
                        if j in own_scene:
                            server.remove_scene(own_scene.pop(j))
            if active == 0 and len(done) == len(jobs):
                nxt = trace[pending[0]].arrival_s if pending else now + 0.002
                time.sleep(max(0.0, min(0.002 if building else 0.01, nxt - (time.perf_counter() - t0))))
    finally:
        if pool:
            pool.shutdown(wait=True)
    wall = time.perf_counter() - t0
    lat = np.array(sorted(done.values()))
    return {
        "requests": len(trace),
        "completed": len(done),
        "wall_s": wall,
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "mean_ms": float(lat.mean() * 1e3) if len(lat) else None,
        "paths_per_s": paths / wall,
        "jobs_per_s": len(done) / wall,
        "offered_rate": len(trace) / max(trace[-1].arrival_s, 1e-9),
        "soup_builds": len(build_s),
        "soup_build_ms_p50": float(np.percentile(build_s, 50) * 1e3) if build_s else None,
        "soup_build_ms_max": float(max(build_s) * 1e3) if build_s else None,
        "server": server.stats(),
    }
