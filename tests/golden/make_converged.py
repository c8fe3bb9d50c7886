"""Writes tests/golden/box_unbiased_fp64.npz (or, with --estimator 0,
box_reference_fp64.npz): the converged image of the box workload (256 x
256, depth 4) rendered by the CPU oracle in FP64 with the unbiased (or the
reference) estimator, and its per-pixel standard error.

The time-to-RMSE measurement compares GPU renders against the unbiased one,
the image tests compare GPU renders of each estimator against its own.  It
uses its own Philox seed, so its noise is independent of every GPU render
(which use seed 1).  Usage (about 20 minutes on 6 threads for 32768
samples/pixel):

    python tests/golden/make_converged.py [--spp 32768] [--threads 6] [--estimator 1]

The fixture holds data only: the per-pixel mean (float32) and the standard
error of that mean (float32), RGB.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED = 0xC0FFEE


def main():
    import oracle
    from tempestrenderer_amd import workloads
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=32768)
    ap.add_argument("--threads", type=int, default=6)
    ap.add_argument("--frame-spp", type=int, default=256)
    ap.add_argument("--estimator", type=int, default=1, help="1 unbiased, 0 the reference estimator")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    out = args.out or os.path.join(HERE, "box_unbiased_fp64.npz" if args.estimator else "box_reference_fp64.npz")
    d = workloads.box(256, 256, samples_camera=args.frame_spp, max_depth=4)
    d.settings.update(estimator=args.estimator, seed=SEED)
    o = oracle.OracleScene(d.width, d.height, d.vpi)
    d.populate(o, oracle.material)
    o.commitScene()
    s = np.zeros((d.height, d.width, 3))
    q = np.zeros((d.height, d.width, 3))
    frames = args.spp // args.frame_spp
    t0 = time.time()
    for f in range(frames):
        fs, fq, _ = o.renderRows(0, d.height, frame=f, fp64=True, threads=args.threads)
        s += fs
        q += fq
        n = (f + 1) * args.frame_spp
        mean = s / n
        var = np.maximum(q / n - mean * mean, 0.0) * n / (n - 1)
        se = np.sqrt(var / n)
        print("frame %d/%d: %d spp, %.0f s, RMSE of the mean %.2e" % (
            f + 1, frames, n, time.time() - t0, float(np.sqrt(np.mean(se * se)))), flush=True)
        if (f + 1) % 16 == 0 or f + 1 == frames:
            np.savez_compressed(out, mean=mean.astype(np.float32), stderr=se.astype(np.float32),
                                spp=np.int64(n), seed=np.int64(SEED), estimator=np.int64(args.estimator))


if __name__ == "__main__":
    main()
