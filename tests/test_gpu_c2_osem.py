"""The full 256x256x81 / 344x252x837 OSEM workload against the reference's own
OsemSinogram3d run (tests/golden/c2_osem.npz: one iteration of 21 subsets from ones,
run by the reference's CPU path in one thread).

The measured sinogram is rebuilt here from the GPU's projection of the phantom: its
sha256 must equal the reference projection's (the whole 72.5M-bin projection
bit-exact), and the Poisson sample drawn from it must equal the reference run's
input (sha256).  The image after the iteration is compared on 131072 seeded voxels
(relative L2 <= 1e-4, NaN voxels coinciding) and by slice sums; the sensitivity of
subset 0 per voxel (1e-5) on the same voxels."""
import hashlib
import json

import numpy as np
import pytest

from apirl_amd import hip, phantoms
from test_gpu_parity import ctx, dev, host, log, torch  # noqa: F401

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [hip.BP_GATHER, hip.BP_SCATTER, hip.BP_TILE])
def test_c2_osem_one_iteration_vs_reference(torch, ctx, golden, mode):
    log(f"case: 256x256x81 OSEM 21 subsets x 1 vs the reference, mode {mode}")
    f = golden("c2_osem.npz")
    g = json.loads(str(f["geom"]))
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])
    img, _ = phantoms.cylinder_points3d(g["nx"], g["nz"], g["dx"], g["dz"])
    proj = host(p.project(dev(torch, img))).ravel()
    assert hashlib.sha256(proj.tobytes()).hexdigest() == str(f["proj_sha256"]), \
        "the full projection differs from the reference's"
    meas = phantoms.poisson_counts(proj, g["counts"], np.random.default_rng(g["seed"]))
    assert hashlib.sha256(meas.tobytes()).hexdigest() == str(f["meas_sha256"])
    del proj
    S = g["subsets"]
    sens = torch.empty((S,) + p.image_shape(), dtype=torch.float32, device="cuda")
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    out, lik = p.reconstruct(dev(torch, meas.reshape(p.subset_shape())), init, g["iters"],
                             subsets=S, bp_mode=mode, sens_out=sens)
    torch.cuda.synchronize()
    final = host(out)
    idx = f["sample_idx"]
    got = final.ravel()[idx].astype(np.float64)
    ref = f["final_sample"].astype(np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(ref)), "NaN voxels differ"
    assert int(np.isnan(final).sum()) == int(f["final_nan"])
    m = np.isfinite(ref)
    e = np.linalg.norm(got[m] - ref[m]) / np.linalg.norm(ref[m])
    log(f"  image rel L2 on {m.sum()} sampled voxels: {e:.2e}")
    assert e <= 1e-4
    sums = np.nansum(final, axis=(1, 2), dtype=np.float64)
    np.testing.assert_allclose(sums, f["final_slice_sums"], rtol=1e-4)
    s0 = host(sens[0]).ravel()[idx].astype(np.float64)
    r0 = f["sens0_sample"].astype(np.float64)
    scale = np.maximum(np.abs(r0), 1e-6 * np.abs(r0).max())
    assert np.max(np.abs(s0 - r0) / scale) <= 1e-5
    np.testing.assert_allclose(host(sens).sum(axis=(1, 2, 3), dtype=np.float64), f["sens_sums"],
                               rtol=1e-5)
    # Likelihoods: the reference adds 3.45M terms (72.5M for the final one) serially
    # in float, which here rounds by up to a few percent; the device sums in double.
    # The first entry is pinned exactly instead: its terms come from the projection of
    # the initial image (bit-exact with the reference's), summed here in double.
    lik = np.asarray(lik, np.float64)
    est = host(p.project(init, subset=0, nsubsets=S)).astype(np.float64).ravel()
    y = meas.reshape(p.nsino, -1, p.nr)[:, 0::S, :].astype(np.float64).ravel()
    nz = est != 0
    exact = np.sum(y[nz] * np.log(est[nz]) - est[nz])
    assert abs(lik[0] - exact) <= 1e-6 * abs(exact), (lik[0], exact)  # float terms, double sum
    np.testing.assert_allclose(lik, f["lik"].astype(np.float64), rtol=5e-2)
