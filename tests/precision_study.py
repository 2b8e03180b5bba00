"""Precision study on the 256x256x81 / 344x252x837 geometry against the FP64 oracle
(run on a GPU host from the repository root: python tests/precision_study.py > study.md;
--iters N sets the OSEM iterations, default 10).

Hot-spheres phantom (cylinder r = 100 mm, 6 spheres of diameter 10-37 mm at 4:1),
Poisson 1e9 counts (numpy generator, seed 0) on its float projection; modes REFERENCE,
KAHAN, F64_SETUP (FP64 only for the ray / plane alpha setup) and F64 (all double).
  * projection of subset 0/21: 2000 sampled bins against the FP64 oracle;
  * gather backprojection of subset 0/21 of that projection: every voxel against the
    FP64 oracle;
  * image after N OSEM iterations x 21 subsets: every voxel against the FP64 oracle's
    OSEM, run here on the CPU (its projector and backprojector over sinogram chunks in
    threads; ctypes releases the GIL, the oracle keeps no shared state).
The OSEM runs use the MLEM threshold rule per subset (Mlem.cpp:39-42: the update where
s >= min_nz + 1e-4 (max - min_nz), else 0) in every mode and in the oracle: with
OSEM's threshold 0 and >= (OsemSinogram3d.cpp:217-237) a voxel that no LOR of one
subset crosses becomes 0/0 = NaN, and in double four voxels of the image's corner
column (x, y) = (0, 0) fall outside subset 1's support (float rounding gives them a
sliver), so that NaN reaches the whole double image through subset 0's 90-degree LORs.
Prints a markdown table; progress goes to stderr."""
import argparse
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import _oracle as O  # noqa: E402
from apirl_amd import hip, phantoms  # noqa: E402

C2 = dict(nx=256, ny=256, nz=81, dx=2.0445, dz=2.027, nproj=252, nr=344, rings=64,
          rfov=261.696, zfov=164.187, rscanner=400.0, span=11, mrd=60)
MODES = {"FP32 (reference)": hip.PREC_REFERENCE, "FP32 + Kahan": hip.PREC_KAHAN,
         "FP32 + FP64 alpha setup": hip.PREC_F64_SETUP, "FP64": hip.PREC_F64}
SUBSETS = 21


def progress(msg):
    print(f"[study] {msg}", file=sys.stderr, flush=True)


def errors(a, ref):
    a = np.asarray(a, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    m = np.isfinite(a) & np.isfinite(ref)
    d = a[m] - ref[m]
    return np.linalg.norm(d) / np.linalg.norm(ref[m]), np.abs(d).max()


class Fp64Osem:
    """The FP64 oracle's OSEM (orc_osem3d_f64's loop, OsemSinogram3d.cpp:24-346) with
    MLEM thresholds, its projections and backprojections split over sinogram chunks."""

    def __init__(self, g, threads):
        self.g = g
        self.im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
        self.ang, self.r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
        self.za, self.zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
        n = len(self.za)
        edges = np.linspace(0, n, 4 * threads + 1).astype(int)
        self.chunks = [(a, b) for a, b in zip(edges[:-1], edges[1:]) if b > a]
        self.pool = ThreadPoolExecutor(threads)

    def angles(self, s):
        return O.subset_angles(self.g["nproj"], SUBSETS, s)

    def project(self, img, ang):
        def one(c):
            a, b = c
            return O.project3d_f64(self.im, img, ang, self.r, self.g["rscanner"],
                                   self.za[a:b], self.zb[a:b])
        return np.concatenate(list(self.pool.map(one, self.chunks)), axis=0)

    def backproject(self, sino, ang):
        def one(c):
            a, b = c
            return O.backproject3d_f64(self.im, sino[a:b], ang, self.r, self.g["rscanner"],
                                       self.za[a:b], self.zb[a:b])
        out = np.zeros((self.im.nz, self.im.ny, self.im.nx), np.float64)
        for part in self.pool.map(one, self.chunks):
            out += part
        return out

    def osem(self, meas, init, iters):
        sens, thr = [], []
        for s in range(SUBSETS):
            idx = self.angles(s)
            ones = np.ones((len(self.za), len(idx), self.g["nr"]), np.float64)
            sn = self.backproject(ones, np.ascontiguousarray(self.ang[idx]))
            nz = sn[sn != 0]
            lo = nz.min() if nz.size else 0.0
            sens.append(sn)
            thr.append(lo + (sn.max() - lo) * 1e-4)
        progress("FP64 oracle: sensitivities done")
        x = np.asarray(init, np.float64).copy()
        for t in range(iters):
            for s in range(SUBSETS):
                idx = self.angles(s)
                ang = np.ascontiguousarray(self.ang[idx])
                y = np.asarray(meas[:, idx, :], np.float64)
                est = self.project(x, ang)
                ratio = np.where((y != 0) & (est != 0), y / np.where(est != 0, est, 1.0), 0.0)
                bp = self.backproject(ratio, ang)
                ok = sens[s] >= thr[s]
                x = np.where(ok, x * bp / np.where(ok, sens[s], 1.0), 0.0)
            progress(f"FP64 oracle: iteration {t + 1}/{iters} done")
        return x


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--threads", type=int, default=16)
    args = ap.parse_args()
    import torch
    g = C2
    ctx = hip.Context(0)
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])
    img = phantoms.cylinder_spheres3d(g["nx"], g["nz"], g["dx"], g["dz"])
    x = torch.from_numpy(img).cuda()
    f64 = Fp64Osem(g, args.threads)
    im, ang, r, za, zb = f64.im, f64.ang, f64.r, f64.za, f64.zb
    idx = f64.angles(0)
    rows = {k: {} for k in MODES}

    progress("projection of subset 0")
    proj = {k: p.project(x, subset=0, nsubsets=SUBSETS, precision=m).cpu().numpy()
            for k, m in MODES.items()}
    rng = np.random.default_rng(0)
    n = 2000
    s, ka, q = (rng.integers(0, len(za), n), rng.integers(0, len(idx), n),
                rng.integers(0, g["nr"], n))
    ref = O.project3d_bins_f64(im, img, ang, r, g["rscanner"], za, zb, s, idx[ka], q)
    for k in MODES:
        rows[k]["proj"] = errors(proj[k][s, ka, q], ref)

    progress("backprojection of subset 0 (FP64 oracle on the CPU)")
    sino = proj["FP32 (reference)"]
    ref = f64.backproject(np.asarray(sino, np.float64), np.ascontiguousarray(ang[idx]))
    st = torch.from_numpy(sino).cuda()
    for k, m in MODES.items():
        b = p.backproject(st, subset=0, nsubsets=SUBSETS, mode=hip.BP_GATHER, precision=m)
        rows[k]["bp"] = errors(b.cpu().numpy(), ref)

    progress(f"{args.iters} OSEM iterations x {SUBSETS} subsets per mode")
    truth = p.project(x).cpu().numpy()
    meas = phantoms.poisson_counts(truth, 1e9, np.random.default_rng(0))
    init = np.ones((g["nz"], g["ny"], g["nx"]), np.float32)
    final = {}
    md, x0 = torch.from_numpy(meas).cuda(), torch.from_numpy(init).cuda()
    for k, m in MODES.items():
        t0 = time.time()
        out, _ = p.reconstruct(md, x0, args.iters, subsets=SUBSETS, bp_mode=hip.BP_GATHER,
                               precision=m, osem_rules=0)
        torch.cuda.synchronize()
        final[k] = out.cpu().numpy()
        progress(f"{k}: {time.time() - t0:.1f} s")
    t0 = time.time()
    ref = f64.osem(meas, init, args.iters)
    progress(f"FP64 oracle OSEM: {time.time() - t0:.0f} s")
    for k in MODES:
        rows[k]["img"] = errors(final[k], ref)

    print(f"| mode | projection rel L2 / max abs | backprojection rel L2 / max abs | "
          f"image after {args.iters} it x {SUBSETS} subsets rel L2 / max abs |")
    print("|---|---|---|---|")
    for k in MODES:
        c = rows[k]
        print(f"| {k} | {c['proj'][0]:.2e} / {c['proj'][1]:.2e} | {c['bp'][0]:.2e} / "
              f"{c['bp'][1]:.2e} | {c['img'][0]:.2e} / {c['img'][1]:.2e} |")


if __name__ == "__main__":
    main()
