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


def bytes_per_closest_ray(pt, scene, cfg, opts_ambient, sample_rows=48, x_step=16, spp=2):
    """Algorithmic HBM bytes per closest-hit ray: 32 (ray) + 16 (hit) +
    128 per 4-wide node + 48 per triangle, with node/triangle counts from
    the oracle's traversal of the exported device BVH over the oracle's own
    path rays (same samples as the device)."""
    import oracle as O
    nodes, tris = pt.bvh_export()
    os_ = O.OracleScene(scene)
    W, H = cfg["width"], cfg["height"]
    rays = []
    for y in np.linspace(0, H - 1, sample_rows).astype(int):
        opts = O.options(W, H, max_bounces=cfg["bounces"], ambient=opts_ambient)
        rays.append(os_.record_rays(W, H, 0, spp, int(y), int(y) + 1, x_step=x_step, max_rays=1 << 16, opts=opts))
    rays = np.concatenate(rays)
    c = O.bvh4_count(nodes, tris, rays).astype(np.float64)
    n_node, n_tri = c[:, 0].mean(), c[:, 1].mean()
    return 32.0 + 16.0 + 128.0 * n_node + 48.0 * n_tri, n_node, n_tri, len(rays)


def cpu_baseline(scene, cfg, ambient, target_s=12.0):
    """The oracle (C++ restatement of the reference's band-parallel
    refine) on a bounded band of rows of the same frame."""
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    os_ = O.OracleScene(scene)
    W, H = cfg["width"], cfg["height"]
    rows = min(H, 4 * threads * 2)
    y0 = max(0, H // 2 - rows // 2)
    opts = O.options(W, H, max_bounces=cfg["bounces"], ambient=ambient)
    t = time.perf_counter()
    os_.render(W, H, 0, 1, opts, y0=y0, y1=y0 + rows, threads=threads)
    dt1 = time.perf_counter() - t
    spp = int(max(1, min(512, target_s / max(dt1, 1e-3))))
    t = time.perf_counter()
    os_.render(W, H, 1, spp, opts, y0=y0, y1=y0 + rows, threads=threads)
    dt = time.perf_counter() - t
    paths = rows * W * spp
    return {"value": paths / dt, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"oracle refine({spp}) of rows {y0}..{y0 + rows} of the {W}x{H} frame "
                      f"({paths} paths, {dt:.1f} s, band-parallel f64 accumulate)"}


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
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {a.gpus}); "
              "the path shards by sample range, see DESIGN.md", file=sys.stderr)
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
    for _ in range(a.warmup):
        pt.refine(a.spp_per_step)
    pt.reset()
    pt.set_timing(True)
    tot = dict(paths=0, closest_rays=0, shadow_rays=0, ms_closest=0.0, ms_anyhit=0.0, ms_shade=0.0, ms_other=0.0,
               launches_closest=0, launches_anyhit=0)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        pt.refine(a.spp_per_step)
        s = pt.stats()
        for k in tot:
            tot[k] += s[k]
    t1 = time.perf_counter()
    dt = t1 - t0
    paths = W * H * a.spp_per_step * a.steps
    value = paths / dt
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
    }
    if not a.no_roofline and tot["launches_closest"]:
        bpr, n_node, n_tri, nr = bytes_per_closest_ray(pt, scene, cfg, ambient)
        avg_ms = tot["ms_closest"] / tot["launches_closest"]
        bytes_per_launch = bpr * tot["closest_rays"] / tot["launches_closest"]
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        traffic = None
        prof = os.path.join(ROOT, "profiles", f"pathtrace_{a.workload}_{W}.json")
        if os.path.exists(prof):
            traffic = json.load(open(prof)).get("hbm_bytes_per_dispatch")
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                           "algorithmic_bytes_per_launch": bytes_per_launch,
                           "kernel": "k_closest<true>", "avg_launch_ms": avg_ms,
                           "bytes_per_ray": bpr, "nodes_per_ray": n_node, "tris_per_ray": n_tri,
                           "rays_sampled": nr}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(scene, cfg, ambient)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
