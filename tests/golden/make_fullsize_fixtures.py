"""Writes the full-size fixtures of the heightfield and volume workloads.

heightfield_ids.npz - the closest-hit contract at full size (1024^2 grid,
    2,093,122 primitives): every pixel-centre primary ray of the 1920x1080
    camera plus 2^20 secondary rays from points on the heightfield
    (workloads.surface_rays).  The ray set is rebuilt from its seeds, so the
    fixture holds its SHA-256, the SHA-256 of the oracle's (geom, prim,
    t bits) arrays, and every 64th expected record.
heightfield_subset_fp64.npz / volume_subset_fp64.npz - converged FP64
    renders (reference estimator, the CPU oracle, Philox seed 0xC0FFEE) of a
    fixed subset of 4096 pixels spread over the full frame, with the
    standard error of each pixel's mean: the yardstick against which a GPU
    render's error is stated.

    python tests/golden/make_fullsize_fixtures.py [--spp-heightfield 8192] [--spp-volume 4096] [--threads 8]

The fixtures hold data only.
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED = 0xC0FFEE
SUBSET = 4096
STRIDE = 64


def id_rays(desc):
    from tempestrenderer_amd import workloads
    return np.concatenate([workloads.primary_rays(desc), workloads.surface_rays(desc, 1 << 20)])


def subset_pixels(desc):
    return np.linspace(0, desc.pixels - 1, SUBSET).astype(np.uint32)


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def hit_digest(hits):
    return digest(hits["geom"], hits["prim"], hits["t"].view(np.uint32))


def converged_subset(desc, spp, frame_spp, threads, out):
    import oracle
    desc.settings.update(seed=SEED, samples_camera=frame_spp)
    o = oracle.OracleScene(desc.width, desc.height, desc.vpi)
    desc.populate(o, oracle.material)
    o.commitScene()
    px = subset_pixels(desc)
    s = np.zeros((px.size, 3))
    q = np.zeros((px.size, 3))
    t0 = time.time()
    frames = spp // frame_spp
    for f in range(frames):
        fs, fq, _ = o.renderPixels(px, frame=f, fp64=True, threads=threads)
        s += fs
        q += fq
        if (f + 1) % 8 == 0 or f + 1 == frames:
            n = (f + 1) * frame_spp
            mean = s / n
            var = np.maximum(q / n - mean * mean, 0.0) * n / (n - 1)
            se = np.sqrt(var / n)
            print("%s: %d spp, %.0f s, rms stderr %.2e" % (os.path.basename(out), n, time.time() - t0,
                                                           float(np.sqrt(np.mean(se * se)))), flush=True)
            np.savez_compressed(out, pixels=px, mean=mean.astype(np.float32), stderr=se.astype(np.float32),
                                spp=np.int64(n), seed=np.int64(SEED))


def main():
    import oracle
    from tempestrenderer_amd import workloads
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp-heightfield", type=int, default=8192)
    ap.add_argument("--spp-volume", type=int, default=4096)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--only", choices=["ids", "heightfield", "volume"], default=None)
    args = ap.parse_args()
    if args.only in (None, "ids"):
        d = workloads.heightfield()
        o = oracle.OracleScene(d.width, d.height, d.vpi)
        d.populate(o, oracle.material)
        o.commitScene()
        rays = id_rays(d)
        hits = o.traceRays(rays, threads=args.threads)
        idx = np.arange(0, rays.shape[0], STRIDE)
        np.savez_compressed(os.path.join(HERE, "heightfield_ids.npz"), n_rays=np.int64(rays.shape[0]),
                            rays_sha256=digest(rays), hits_sha256=hit_digest(hits), stride=np.int64(STRIDE),
                            geom=hits["geom"][idx], prim=hits["prim"][idx], t=hits["t"][idx])
        print("ids: %d rays, %.1f%% hit" % (rays.shape[0], 100.0 * np.mean(hits["geom"] != 0xFFFFFFFF)))
    if args.only in (None, "heightfield"):
        converged_subset(workloads.heightfield(), args.spp_heightfield, 256, args.threads,
                         os.path.join(HERE, "heightfield_subset_fp64.npz"))
    if args.only in (None, "volume"):
        converged_subset(workloads.volume(), args.spp_volume, 256, args.threads,
                         os.path.join(HERE, "volume_subset_fp64.npz"))


if __name__ == "__main__":
    main()
