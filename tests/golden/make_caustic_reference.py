"""Writes the converged reference of a window of the glossy-caustic frame
that bench.py's time-to-MSE leg measures against: plain sampling (no
guide) through the HIP path tracer at a fixed seed that the bench does not
use, accumulated in f64, stored as the per-pixel mean and the variance of
that mean (from the sums of squares).

    python tests/golden/make_caustic_reference.py --window X0 Y0 W H \\
        --spp N --out tests/golden/caustic_window_reference.npz

Needs a HIP device. tests/test_oracle_guide_reference.py checks the stored
reference against the oracle on a crop, within 3 sigma."""
import argparse
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oxideav-render_amd")]

import numpy as np  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--window", type=int, nargs=4, required=True, metavar=("X0", "Y0", "W", "H"))
    ap.add_argument("--spp", type=int, required=True)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--frame", type=int, default=1024)
    ap.add_argument("--bounces", type=int, default=8)
    ap.add_argument("--pass-spp", type=int, default=4096)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    import oxr
    from oxr import scenes
    sc = scenes.glossy_caustic()
    pt = oxr.PathTracer(0)
    o = oxr.RenderOptions(width=a.frame, height=a.frame, ambient=sc.ambient)
    o.path_trace.samples_per_pixel = a.spp
    o.path_trace.max_bounces = a.bounces
    o.path_trace.seed = a.seed
    pt.sync(sc, o)
    x0, y0, w, h = a.window
    pt.set_window(x0, y0, w, h)
    t = time.time()
    done = 0
    while done < a.spp:
        k = min(a.pass_spp, a.spp - done)
        pt.refine(k)
        done += k
        if done % (128 * a.pass_spp) == 0 or done == a.spp:
            print(f"{done} / {a.spp} spp, {time.time() - t:.1f} s", flush=True)
    acc, sq = pt.accum(with_sum_sq=True)
    pt.close()
    n = acc[..., 3:4]
    if not (n == a.spp).all():
        raise SystemExit("the window has uncovered samples; choose one inside the room")
    mean = acc[..., :3] / n
    var_mean = np.maximum(sq / n - mean * mean, 0.0) / n
    np.savez_compressed(a.out, mean=mean, var_mean=var_mean, window=np.array(a.window), spp=a.spp, seed=a.seed,
                        frame=np.array([a.frame, a.frame]), max_bounces=a.bounces)
    print(f"wrote {a.out}: mean {mean.mean():.6f}, mean variance of the mean {var_mean.mean():.3e}", flush=True)


if __name__ == "__main__":
    main()
