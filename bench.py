#!/usr/bin/env python3
"""Throughput bench of the MI355X PathTrace backend.

Workload (default): the material sweep — 7x7 UV spheres (64x32 segments)
on a checker floor, point + spot light, sky 0.3; 1024x1024, 8 bounces,
Owen-Sobol' seed 0, f32 radiance with f64 accumulation. One step = one
`PathTracer::refine(spp_per_step)` over the whole frame, scene resident in
HBM; `--steps 16 --spp-per-step 16` is one full 256-spp render.

Prints one JSON line: paths/s (value), rays/s, the roofline of the
dominant kernel (closest-hit traversal: algorithmic bytes per launch /
HIP-event launch time, against 8 TB/s HBM) and the oracle's CPU
restatement timed on this host's cores on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "oxideav-render_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

METRIC = json.load(open(os.path.join(ROOT, "BASELINE.json")))["metric"] if os.path.exists(
    os.path.join(ROOT, "BASELINE.json")) else "paths/s"
HBM_PEAK_GBS = 8000.0


def workload(name):
    from oxr import scenes
    if name == "material_sweep":
        return scenes.material_sweep(), dict(width=1024, height=1024, spp=256, bounces=8)
    if name == "cornell":
        return scenes.cornell(), dict(width=256, height=256, spp=64, bounces=8)
    if name == "glossy_caustic":
        return scenes.glossy_caustic(), dict(width=1024, height=1024, spp=64, bounces=8)
    if name == "terrain":
        return scenes.terrain(), dict(width=1920, height=1080, spp=1024, bounces=8)
    raise SystemExit(f"unknown workload {name}")


def traversal_counts(pt, spp):
    """Node and triangle fetches per ray, counted by the device itself (the
    counting closest-hit and occlusion variants) over every ray of one
    refine(spp) of the workload; the accumulation is reset afterwards.
    Algorithmic HBM bytes per closest ray: 32 (ray) + 16 (hit) + 64 per
    4-wide node (quantised child boxes) + 48 per triangle; per occlusion
    ray: 32 (ray) + 4 (slot) + 1 (visibility) + 64 per node + 48 per
    triangle tested."""
    pt.set_timing(False, count=True)
    pt.refine(spp)
    s = pt.stats()
    pt.set_timing(False)
    pt.reset()
    nc, ns = max(s["closest_rays"], 1), max(s["shadow_rays"], 1)
    c_node, c_tri = s["closest_nodes"] / nc, s["closest_tris"] / nc
    a_node, a_tri = s["anyhit_nodes"] / ns, s["anyhit_tris"] / ns
    return {"closest": (32.0 + 16.0 + 64.0 * c_node + 48.0 * c_tri, c_node, c_tri, s["closest_rays"]),
            "anyhit": (32.0 + 4.0 + 1.0 + 64.0 * a_node + 48.0 * a_tri, a_node, a_tri, s["shadow_rays"])}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def lease_threads():
    """Host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where the host states its share that way."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(n, 1)


def converged_error(acc_g, sq_g, n_g, acc_c, sq_c, n_c):
    """GPU mean (n_g samples, seed 0) against the oracle's f64 mean (n_c
    samples, another seed) over the same pixels, per pixel and channel:
    z = |m_g - m_c| / sqrt(var / n_g + var / n_c) with var the per-sample
    variance pooled over both runs, on pixels every sample of both covered."""
    full = (acc_g[..., 3] == n_g) & (acc_c[..., 3] == n_c)
    mg, mc = acc_g[..., :3][full] / n_g, acc_c[..., :3][full] / n_c
    # One estimator, two seeds: the per-sample variance pooled over both.
    n = n_g + n_c
    var = np.maximum((sq_g[full] + sq_c[full]) / n - ((acc_g[..., :3][full] + acc_c[..., :3][full]) / n) ** 2, 0.0)
    sig = np.sqrt(var / n_g + var / n_c)
    d = np.abs(mg - mc)
    z = np.where(sig > 0, d / np.where(sig > 0, sig, 1.0), np.where(d > 0, np.inf, 0.0))
    return {"pixels": int(full.sum()), "gpu_spp": int(n_g), "cpu_spp": int(n_c),
            "max_z": float(z.max()) if z.size else None,
            "frac_gt_3sigma": float((z > 3).mean()) if z.size else None,
            "mean_gpu": float(mg.mean()) if mg.size else None, "mean_cpu": float(mc.mean()) if mc.size else None}


def cpu_baseline(scene, cfg, ambient, gpu_accum=None, target_s=10.0, single_s=5.0, runs=3):
    """The oracle (C++ restatement of the reference's band-parallel
    refine) on a centred band of rows of the same frame, with every core of
    the lease: the median of `runs` timed runs of at least `target_s`, then
    one run on one thread over 4 rows. The oracle samples with seed 1, so
    its last run is also an estimate independent of the GPU's: with
    gpu_accum = (accum, sum_sq, spp) of the GPU frame (seed 0) the per-pixel
    error over the band is reported (converged_error)."""
    import oracle as O
    threads = lease_threads()
    os_ = O.OracleScene(scene)
    W, H = cfg["width"], cfg["height"]
    opts = O.options(W, H, max_bounces=cfg["bounces"], ambient=ambient, seed=1)

    def timed(rows, nthreads, budget_s, nruns):
        y0 = max(0, H // 2 - rows // 2)
        t = time.perf_counter()
        os_.render(W, H, 0, 1, opts, y0=y0, y1=y0 + rows, threads=nthreads)
        dt1 = time.perf_counter() - t
        spp = int(max(1, min(1024, math.ceil(budget_s / max(dt1, 1e-3)))))
        rates, last = [], None
        for _ in range(nruns):
            t = time.perf_counter()
            last = os_.render(W, H, 1, spp, opts, y0=y0, y1=y0 + rows, threads=nthreads, sum_sq=True)
            rates.append(rows * W * spp / (time.perf_counter() - t))
        return float(np.median(rates)), rates, spp, y0, last

    rows = min(H, 8 * threads)
    v, rates, spp, y0, (acc_c, sq_c) = timed(rows, threads, target_s, runs)
    v1, _, spp1, y01, _ = timed(4, 1, single_s, 1)
    out = {"value": v, "unit": "paths/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
           "hardware_threads": os.cpu_count(), "single_thread_paths_per_s": v1,
           "runs_paths_per_s": rates,
           "sample": f"oracle refine({spp}) of rows {y0}..{y0 + rows} of the {W}x{H} frame "
                     f"({rows * W * spp} paths per run, median of {runs} runs of >= {target_s:.0f} s, "
                     f"{threads} threads, band-parallel f64 accumulate, seed 1); one thread: "
                     f"refine({spp1}) of rows {y01}..{y01 + 4} (one run)"}
    if gpu_accum is not None:
        ag, sg, n_g = gpu_accum
        err = converged_error(ag[y0:y0 + rows], sg[y0:y0 + rows], n_g, acc_c, sq_c, spp)
        err["rows"] = [y0, y0 + rows]
        out["error_vs_gpu"] = err
    return out


def serving_bench(n=512, lam=40.0, seed=0, rates=(32.0, 28.0, 24.0, 20.0, 16.0, 12.0, 8.0), n_rate=160):
    """The serving mix as configured: n requests with Poisson arrivals at
    rate lam, each drawing the Cornell box, the material sweep or a fresh
    seeded triangle soup (N log-uniform 1k..2M, generated and prepared on
    build threads when the request arrives, its build counted in the
    latency, freed with the job), resolution in {128..1024}^2, spp in
    {4, 16, 64, 256}, 4..8 bounces. Reports throughput and p50/p99
    latency, then the same mix at lower rates until one is sustained
    (completions keep up with arrivals): the highest sustained rate found."""
    from oxr import scenes
    from oxr.serve import Server, make_trace, run_requests
    sc = [scenes.cornell(), scenes.material_sweep()]
    srv = Server(0)
    t = time.perf_counter()
    ids = [srv.add_scene(x) for x in sc]
    setup_s = time.perf_counter() - t

    def one(n_req, rate, sd):
        trace = make_trace(n=n_req, lam=rate, n_scenes=3, seed=sd, soups=True)
        out = run_requests(srv, ids, sc, trace)
        out["sustained"] = out["jobs_per_s"] >= 0.95 * out["offered_rate"]
        paths = [r.width * r.height * r.spp for r in trace]
        out["offered_paths_per_s"] = float(sum(paths)) / max(trace[-1].arrival_s, 1e-9)
        out["lambda"] = rate
        return out

    out = one(n, lam, seed)
    out["scene_setup_s"] = setup_s
    if not out["sustained"]:
        tried = []
        for k, rate in enumerate(rates):
            if rate >= lam:
                continue
            o = one(n_rate, rate, seed + 1 + k)
            o.pop("server", None)
            tried.append({key: o[key] for key in ("lambda", "requests", "jobs_per_s", "offered_rate", "p50_ms",
                                                   "p99_ms", "paths_per_s", "offered_paths_per_s", "sustained")})
            if o["sustained"]:
                out["highest_sustained"] = tried[-1]
                break
        out["lower_rates"] = tried
    srv.close()
    return out


def roofline_of(pt, spp, tot, steps, workload, W):
    """Rooflines of the two traversal kernels over one refine(spp) step:
    algorithmic bytes per launch (device-counted fetches) over the average
    HIP-event launch time, against 8 TB/s. `traffic` is the measured HBM
    bytes per launch (the copy-calibrated FETCH_SIZE + WRITE_SIZE per ray in
    profiles/, times this run's rays per launch); where it is well below
    the algorithmic bytes the BVH is served from L2 / MALL and `achieved` is
    an effective bandwidth, not HBM's."""
    counts = traversal_counts(pt, spp)
    prof = os.path.join(ROOT, "profiles", f"pathtrace_{workload}_{W}.json")
    pj = json.load(open(prof)) if os.path.exists(prof) else {}
    out = {}
    for kern, rays_key, ms_key, n_key, pkey in (
            ("closest", "closest_rays", "ms_closest", "launches_closest", "hbm_bytes_per_closest_ray"),
            ("anyhit", "shadow_rays", "ms_anyhit", "launches_anyhit", "hbm_bytes_per_shadow_ray")):
        if not tot.get(n_key):
            continue
        bpr, n_node, n_tri, nr = counts[kern]
        avg_ms = tot[ms_key] / tot[n_key]
        bytes_per_launch = bpr * tot[rays_key] / tot[n_key]
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        traffic = pj[pkey] * tot[rays_key] / tot[n_key] if pkey in pj else None
        r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
             "algorithmic_bytes_per_launch": bytes_per_launch,
             "kernel": "k_closest<true>" if kern == "closest" else "k_anyhit<true>", "avg_launch_ms": avg_ms,
             "bytes_per_ray": bpr, "nodes_per_ray": n_node, "tris_per_ray": n_tri, "rays_counted": nr}
        if traffic is not None:
            r["hbm_traffic_gbs"] = traffic / (avg_ms * 1e-3) / 1e9
            r["hbm_traffic_frac"] = r["hbm_traffic_gbs"] / HBM_PEAK_GBS
            r["kind"] = ("effective: the BVH is served from L2 / MALL (measured HBM traffic is "
                         f"{traffic / bytes_per_launch:.2f}x the algorithmic bytes)"
                         if traffic < 0.5 * bytes_per_launch else "HBM")
        if kern in pj.get("counters", {}):
            r["counters"] = pj["counters"][kern]
        out[kern] = r
    return out


def terrain_leg(spp_per_step=4, steps=2, crop=16, cpu_spp_factor=16, error_spp=1024, autotune=True):
    """The large-scene config at its full size (12.4M triangles, 1920x1080,
    8 bounces), bounded to a few steps of spp_per_step: paths/s, rays/s,
    the rooflines of both traversal kernels from the device's own fetch
    counts, and the error against the oracle: the GPU frame at the config's
    error_spp samples (seed 0) against the oracle's f64 accumulation of a
    centred crop at cpu_spp_factor x the samples (seed 1). The terrain's
    small bright emitters make single samples heavy-tailed; at the config's
    1024 spp the 3-sigma count behaves as for a normal difference. The
    launch configuration is autotuned first, as for the headline workload."""
    import oxr
    scene, cfg = workload("terrain")
    W, H = cfg["width"], cfg["height"]
    pt = oxr.PathTracer(0)
    opts = oxr.RenderOptions(width=W, height=H, ambient=getattr(scene, "ambient", 0.0),
                             path_trace=oxr.PathTraceOptions(samples_per_pixel=cfg["spp"], max_bounces=cfg["bounces"]))
    t = time.perf_counter()
    pt.sync(scene, opts)
    sync_s = time.perf_counter() - t
    tune = pt.autotune(spp=spp_per_step) if autotune else None
    pt.refine(spp_per_step)
    pt.reset()
    pt.set_timing(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        pt.refine(spp_per_step)
    dt = time.perf_counter() - t0
    pt.set_timing(True)
    pt.refine(spp_per_step)
    st = pt.stats()
    pt.set_timing(False)
    tot = {k: st[k] * steps for k in ("closest_rays", "shadow_rays", "ms_closest", "ms_anyhit", "ms_shade",
                                      "ms_other")}
    tot["launches_closest"] = st["launches_closest"] * steps
    tot["launches_anyhit"] = st["launches_anyhit"] * steps
    out = {"workload": f"terrain {W}x{H}, {scene.num_tris} tris, {cfg['bounces']} bounces, "
                       f"{steps * spp_per_step} spp timed",
           "paths_per_s": W * H * spp_per_step * steps / dt,
           "rays_per_s": (tot["closest_rays"] + tot["shadow_rays"]) / dt, "ms_per_step": dt * 1e3 / steps,
           "kernel_ms": {"closest": tot["ms_closest"], "anyhit": tot["ms_anyhit"], "shade": tot["ms_shade"],
                         "other": tot["ms_other"]},
           "scene_sync_s": sync_s, "autotune": tune, "bvh": pt.bvh_info()}
    if tot["launches_closest"]:
        rl = roofline_of(pt, spp_per_step, tot, steps, "terrain", W)
        out["roofline"] = rl["closest"]
        if "anyhit" in rl:
            out["roofline_anyhit"] = rl["anyhit"]
    # The error leg: the frame at error_spp (roofline_of reset the tracer).
    for _ in range(error_spp // 16):
        pt.refine(16)
    acc_g, sq_g = pt.accum(with_sum_sq=True)
    pt.close()
    import oracle as O
    t = time.perf_counter()
    os_ = O.OracleScene(scene)
    y0, x0 = H // 2 - crop // 2, W // 2 - crop // 2
    n_g = (error_spp // 16) * 16
    n_c = cpu_spp_factor * n_g
    oo = O.options(W, H, max_bounces=cfg["bounces"], ambient=getattr(scene, "ambient", 0.0), seed=1)
    ac, sc = os_.render_rect(W, H, 0, n_c, x0, x0 + crop, y0, y0 + crop, opts=oo, threads=lease_threads())
    err = converged_error(acc_g[y0:y0 + crop, x0:x0 + crop], sq_g[y0:y0 + crop, x0:x0 + crop], n_g, ac, sc, n_c)
    err["crop"] = [y0, y0 + crop, x0, x0 + crop]
    err["oracle_s"] = time.perf_counter() - t
    out["error"] = err
    return out


# Online training budget of the guided leg: 80 Adam steps of 16384
# samples. On the caustic window, 40-160 steps (16k-64k samples each) all
# reach MSE 1e-3 in 86-92k spp, so the extra steps only add time; at 20
# steps the guide is undertrained and needs 6x the samples of plain.
GUIDE_BENCH_CONFIG = {"max_steps": 80, "batch": 16384}
CAUSTIC_REFERENCE = os.path.join(ROOT, "tests", "golden", "caustic_window_reference.npz")


def window_mse(acc, ref_mean, ref_var_mean):
    """MSE of a window's mean image against the converged reference, per
    pixel and channel, less the reference's own variance of the mean (so
    the figure estimates the error against the exact expectation)."""
    n = acc[..., 3:4]
    m = np.where(n > 0, acc[..., :3] / np.maximum(n, 1.0), 0.0)
    return float(((m - ref_mean) ** 2).mean() - ref_var_mean.mean())


def guide_mse_bench(target=1e-3, pass_spp=256, max_s=40.0, guide_cfg=None, tile=8):
    """Time to MSE `target`, plain vs guided, on a pixel-exact window of the
    1024x1024 glossy-caustic frame: each method refines the window in
    passes of pass_spp until the MSE of its image against the committed
    f64 reference of the window (tests/golden/caustic_window_reference.npz:
    plain sampling, another seed, at least 16x the samples plain needs here)
    falls to `target`. Wall time of the refine calls; the guided time
    includes its online training (train_ms) and inference (infer_ms). The
    two final images are compared by tile means as the unbiasedness check."""
    import oxr
    from oxr import scenes
    from oxr.guide import Guide
    ref = np.load(CAUSTIC_REFERENCE)
    x0, y0, w, h = (int(v) for v in ref["window"])
    W = int(ref["frame"][0])
    sc = scenes.glossy_caustic()
    gcfg = dict(GUIDE_BENCH_CONFIG if guide_cfg is None else guide_cfg)

    def run(guided):
        pt = oxr.PathTracer(0)
        o = oxr.RenderOptions(width=W, height=W, ambient=sc.ambient)
        o.path_trace.samples_per_pixel = 1 << 30
        o.path_trace.max_bounces = int(ref["max_bounces"])
        o.path_trace.seed = 0
        pt.sync(sc, o)
        pt.set_window(x0, y0, w, h)
        g = Guide(0, **gcfg) if guided else None
        if g:
            pt.set_guide(g, train=True)
        pt.refine(pass_spp)  # warm-up: buffers, graph capture
        if g:  # start from an untrained guide: its training is part of the measured time
            pt.set_guide(None)
            g.close()
            g = Guide(0, **gcfg)
            pt.set_guide(g, train=True)
        pt.reset()
        t_tot, curve, hit = 0.0, [], None
        while t_tot < max_s:
            t = time.perf_counter()
            pt.refine(pass_spp)
            t_tot += time.perf_counter() - t
            acc, sq = pt.accum(with_sum_sq=True)
            mse = window_mse(acc, ref["mean"], ref["var_mean"])
            curve.append((pt.samples(), round(t_tot, 4), mse))
            if mse <= target:
                hit = (pt.samples(), t_tot)
                break
        st = g.stats() if g else None
        infer_pass_ms = None
        if g:
            # Inference runs inside the replayed bounce graph: its device
            # time comes from one more pass with per-kernel events.
            pt.set_timing(True)
            pt.refine(pass_spp)
            infer_pass_ms = g.stats()["infer_ms"] - st["infer_ms"]
            pt.set_timing(False)
        pt.close()
        if g:
            g.close()
        n = np.maximum(acc[..., 3:4], 1.0)
        m = acc[..., :3] / n
        var_m = np.maximum(sq / n - m * m, 0.0) / n
        res = {"how": "measured" if hit else "not reached",
               "time_to_mse_s": hit[1] if hit else None, "spp": hit[0] if hit else None,
               "passes": len(curve), "final_mse": curve[-1][2], "final_spp": curve[-1][0],
               "s_per_pass": t_tot / max(len(curve), 1), "curve": curve[:: max(1, len(curve) // 16)],
               "guide": st}
        if st:
            res["train_ms"] = st["train_ms"]  # device time of the Adam steps within the measured passes
            res["infer_ms"] = infer_pass_ms * len(curve)  # one timed pass's inference x the passes
            res["infer_ms_per_pass"] = infer_pass_ms
        return res, m, var_m

    plain, pm, pv = run(False)
    guided, gm, gv = run(True)
    # Unbiasedness: tile means (tile x tile pixels) of the two images.
    th, tw = h // tile, w // tile
    tmean = lambda a: a[:th * tile, :tw * tile].reshape(th, tile, tw, tile, 3).mean((1, 3))  # noqa: E731
    tvar = lambda v: v[:th * tile, :tw * tile].reshape(th, tile, tw, tile, 3).sum((1, 3)) / tile ** 4  # noqa: E731
    z = (tmean(gm) - tmean(pm)) / np.sqrt(tvar(gv) + tvar(pv) + 1e-30)
    out = {"workload": f"glossy_caustic {W}x{W} frame, window {w}x{h} at ({x0}, {y0}), "
                       f"{int(ref['max_bounces'])} bounces, {pass_spp} spp per pass",
           "target_mse": target, "reference": {"spp": int(ref["spp"]), "seed": int(ref["seed"]),
                                               "var_of_mean": float(ref["var_mean"].mean())},
           "guide_config": gcfg, "plain": plain, "guided": guided,
           "unbiased_check": {"tile": tile, "tiles": int(z.size), "max_abs_z": float(np.abs(z).max()),
                              "frac_abs_z_gt_3": float((np.abs(z) > 3).mean()),
                              "mean_plain": float(pm.mean()), "mean_guided": float(gm.mean())}}
    if plain["time_to_mse_s"] and guided["time_to_mse_s"]:
        out["guided_over_plain_time"] = guided["time_to_mse_s"] / plain["time_to_mse_s"]
    return out


def copy_calibration(mib):
    """A plain device-to-device copy of known size (HIP's copy kernel), for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp-per-step", type=int, default=16)
    ap.add_argument("--workload", default="material_sweep")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-guide", action="store_true", help="skip the guided vs plain time-to-MSE leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop latency leg")
    ap.add_argument("--no-terrain", action="store_true", help="skip the 12.4M-triangle terrain leg")
    ap.add_argument("--terrain-only", action="store_true")
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--guide-only", action="store_true")
    ap.add_argument("--guide-pass-spp", type=int, default=256)
    ap.add_argument("--guide-config", default=None, help="JSON overrides of the guide configuration")
    ap.add_argument("--guide-max-s", type=float, default=40.0)
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.serving_only or a.guide_only or a.terrain_only:
        import oxr
        if oxr.device_count() == 0:
            raise SystemExit("bench.py: no HIP device visible")
        print(json.dumps(serving_bench() if a.serving_only else terrain_leg() if a.terrain_only else
                         guide_mse_bench(pass_spp=a.guide_pass_spp, max_s=a.guide_max_s,
                                        guide_cfg=json.loads(a.guide_config) if a.guide_config else None)), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {a.gpus}); "
              "the path shards by sample range across GPUs", file=sys.stderr)
        raise SystemExit(2)

    import oxr
    if oxr.device_count() == 0:
        raise SystemExit("bench.py: no HIP device visible")
    scene, cfg = workload(a.workload)
    W, H = cfg["width"], cfg["height"]
    ambient = getattr(scene, "ambient", 0.0)
    pt = oxr.PathTracer(0)
    opts = oxr.RenderOptions(width=W, height=H, ambient=ambient,
                             path_trace=oxr.PathTraceOptions(samples_per_pixel=cfg["spp"],
                                                             max_bounces=cfg["bounces"]))
    t = time.perf_counter()
    pt.sync(scene, opts)
    sync_s = time.perf_counter() - t
    tune = pt.autotune(spp=min(a.spp_per_step, 16)) if not a.no_autotune else None
    for _ in range(a.warmup):
        pt.refine(a.spp_per_step)
    pt.reset()
    # Timed region: kernels only (no per-kernel events), graphs replayed.
    pt.set_timing(False)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        pt.refine(a.spp_per_step)
    t1 = time.perf_counter()
    dt = t1 - t0
    paths = W * H * a.spp_per_step * a.steps
    value = paths / dt
    gpu_accum = pt.accum(with_sum_sq=True) + (a.spp_per_step * a.steps,) if not a.no_cpu_baseline else None
    # One more step with HIP events around every kernel (eager launches):
    # the per-kernel breakdown and ray counts, scaled to the timed steps.
    pt.set_timing(True)
    pt.refine(a.spp_per_step)
    s = pt.stats()
    tot = {k: s[k] * a.steps for k in ("paths", "closest_rays", "shadow_rays", "ms_closest", "ms_anyhit",
                                       "ms_shade", "ms_other")}
    tot["launches_closest"] = s["launches_closest"] * a.steps
    tot["launches_anyhit"] = s["launches_anyhit"] * a.steps
    rays = tot["closest_rays"] + tot["shadow_rays"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt * 1000.0 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": f"{a.workload} {W}x{H}, {a.spp_per_step * a.steps} spp timed, "
                               f"{cfg['bounces']} bounces, {scene.num_tris} tris",
                   "spp_per_step": a.spp_per_step, "max_bounces": cfg["bounces"], "width": W, "height": H},
        "rays_per_s": rays / dt,
        "closest_rays": tot["closest_rays"],
        "shadow_rays": tot["shadow_rays"],
        "kernel_ms": {"closest": tot["ms_closest"], "anyhit": tot["ms_anyhit"], "shade": tot["ms_shade"],
                      "other": tot["ms_other"]},
        "scene_sync_s": sync_s,
        "bvh": pt.bvh_info(),
        "autotune": tune,
        "kernel_ms_per_step_source": "one extra step with per-kernel HIP events, x steps",
    }
    if not a.no_roofline and tot["launches_closest"]:
        rl = roofline_of(pt, a.spp_per_step, tot, a.steps, a.workload, W)
        out["roofline"] = rl["closest"]
        if "anyhit" in rl:
            out["roofline_anyhit"] = rl["anyhit"]
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, cfg, ambient, gpu_accum=gpu_accum)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        out["error"] = dict(out["cpu_baseline"].pop("error_vs_gpu"),
                            what="per pixel and channel |GPU mean - oracle mean| / sigma of the difference over "
                                 "the CPU leg's rows (oracle seed 1, GPU seed 0: independent estimates)")
    pt.close()
    # The other two parts of the metric, each reported beside the paths/s
    # value (an error in one is reported, not fatal to the line).
    guide_leg = lambda: guide_mse_bench(pass_spp=a.guide_pass_spp, max_s=a.guide_max_s,  # noqa: E731
                                        guide_cfg=json.loads(a.guide_config) if a.guide_config else None)
    legs = (("large_scene", a.no_terrain or a.workload == "terrain",
             lambda: terrain_leg(autotune=not a.no_autotune)),
            ("guided_time_to_mse", a.no_guide, guide_leg),
            ("serving", a.no_serving, serving_bench))
    for key, skip, fn in legs:
        if skip:
            continue
        try:
            out[key] = fn()
        except Exception as e:  # noqa: BLE001
            out[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
