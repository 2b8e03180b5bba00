"""Precision study on the 256x256x81 / 344x252x837 geometry (run on a GPU host from
the repository root: python tests/precision_study.py > study.md).

Hot-spheres phantom (cylinder r = 100 mm, 6 spheres of diameter 10-37 mm at 4:1),
Poisson 1e9 counts (torch generator, seed 0); modes REFERENCE, KAHAN, F64_ALPHA.
  * projection of subset 0/21: 2000 sampled bins against the FP64 oracle;
  * gather backprojection of subset 0/21 of that projection: every voxel against the
    FP64 oracle (a few seconds of single-thread CPU);
  * image after 10 OSEM iterations x 21 subsets: against the F64_ALPHA image, which
    stands in for the FP64 oracle here (the oracle's OSEM would take hours of CPU at
    this size; on the 64x64x17 case of test_gpu_precision.py F64_ALPHA is within
    ~4e-7 of it).  The update uses the MLEM threshold rule (Mlem.cpp:39-42) in every
    mode: with OSEM's threshold 0 and >= (OsemSinogram3d.cpp:217-237) a voxel that no
    LOR of a subset crosses becomes 0/0 = NaN, and in double four voxels of the
    image's corner column (x, y) = (0, 0) fall outside subset 1's support (float
    rounding gives them a sliver), so that NaN reaches the whole image through
    subset 0's 90-degree LORs.
Prints a markdown table."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import _oracle as O  # noqa: E402
from apirl_amd import hip, phantoms  # noqa: E402

C2 = dict(nx=256, ny=256, nz=81, dx=2.0445, dz=2.027, nproj=252, nr=344, rings=64,
          rfov=261.696, zfov=164.187, rscanner=400.0, span=11, mrd=60)
MODES = {"FP32 (reference)": hip.PREC_REFERENCE, "FP32 + Kahan": hip.PREC_KAHAN,
         "FP32 + FP64 alpha setup": hip.PREC_F64_ALPHA}


def progress(msg):
    print(f"[study] {msg}", file=sys.stderr, flush=True)


def errors(a, ref):
    a = np.asarray(a, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    m = np.isfinite(a) & np.isfinite(ref)
    d = a[m] - ref[m]
    return np.linalg.norm(d) / np.linalg.norm(ref[m]), np.abs(d).max()


def main():
    import torch
    g = C2
    ctx = hip.Context(0)
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])
    img = phantoms.cylinder_spheres3d(g["nx"], g["nz"], g["dx"], g["dz"])
    x = torch.from_numpy(img).cuda()
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    ang, r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    idx = O.subset_angles(g["nproj"], 21, 0)
    rows = {k: {} for k in MODES}

    progress("projection of subset 0")
    proj = {k: p.project(x, subset=0, nsubsets=21, precision=m).cpu().numpy()
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
    t0 = time.time()
    ref = O.backproject3d_f64(im, sino, ang[idx], r, g["rscanner"], za, zb)
    progress(f"oracle backprojection took {time.time() - t0:.0f} s")
    st = torch.from_numpy(sino).cuda()
    for k, m in MODES.items():
        b = p.backproject(st, subset=0, nsubsets=21, mode=hip.BP_GATHER, precision=m)
        rows[k]["bp"] = errors(b.cpu().numpy(), ref)

    progress("10 OSEM iterations x 21 subsets per mode")
    truth = p.project(x)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0)
    meas = torch.poisson(truth * (1e9 / float(truth.double().sum())), generator=gen)
    init = torch.ones_like(x)
    final = {}
    for k, m in MODES.items():
        t0 = time.time()
        out, _ = p.reconstruct(meas, init, 10, subsets=21, bp_mode=hip.BP_GATHER, precision=m,
                               osem_rules=0)
        torch.cuda.synchronize()
        final[k] = out.cpu().numpy()
        progress(f"{k}: {time.time() - t0:.1f} s")
    for k in MODES:
        rows[k]["img"] = errors(final[k], final["FP32 + FP64 alpha setup"])

    print("| mode | projection rel L2 / max abs | backprojection rel L2 / max abs | "
          "OSEM image vs F64_ALPHA rel L2 / max abs |")
    print("|---|---|---|---|")
    for k in MODES:
        c = rows[k]
        print(f"| {k} | {c['proj'][0]:.2e} / {c['proj'][1]:.2e} | {c['bp'][0]:.2e} / "
              f"{c['bp'][1]:.2e} | {c['img'][0]:.2e} / {c['img'][1]:.2e} |")


if __name__ == "__main__":
    main()
