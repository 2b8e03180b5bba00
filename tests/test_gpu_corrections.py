"""Corrections and mMR subsets on the GPU against the reference's goldens
(tests/golden/corrections.npz, written by make_goldens.py from the reference's own
OsemSinogram3d / MlemSinogram3d runs).

* OSEM with an additive sinogram varying by angle and a multiplicative sinogram
  (OsemSinogram3d.cpp:101-117, 172-173, 327-346): the reference adds the additive bin
  at the subset-local angle index (Sinogram3D::addBinToBin on a subset), without a
  clamp; the multiplicative sinogram enters the subset sensitivities only.
* Additive bins below minus the projection: no clamp, so the likelihood is NaN
  wherever the reference's is.
* OSEM on the mMR with 3 subsets: the subsets trace cylindrical LORs on 328 mm
  (APIRL_SCANNER_MMR_SUBSET), the final projection the mMR's.
* MLEM3D on the mMR with both terms.

Bars: sensitivities within 1e-5 per voxel, images within 1e-4 relative L2 (NaN voxels
coinciding), the subset projection bit-exact, finite likelihoods within 1e-4."""
import hashlib
import json
import sys

import numpy as np
import pytest

from apirl_amd import hip, phantoms

pytestmark = pytest.mark.gpu

BP_RTOL = 1e-5
BP_FLOOR = 1e-6
IMG_RL2 = 1e-4


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    return t


@pytest.fixture(scope="module")
def ctx(torch):
    return hip.Context(0)


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    m = np.isfinite(a) & np.isfinite(b)
    return np.linalg.norm(a[m] - b[m]) / np.linalg.norm(b[m])


def assert_voxel_close(a, b, rtol=BP_RTOL, floor=BP_FLOOR):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    scale = np.maximum(np.abs(b), floor * np.abs(b).max())
    err = np.abs(a - b) / scale
    worst = int(np.argmax(err))
    assert err.max() <= rtol, f"max rel err {err.max():.3e} at {worst}: {a[worst]} vs {b[worst]}"


def assert_image(out, ref):
    assert np.array_equal(np.isnan(out), np.isnan(ref)), "NaN positions differ"
    e = rel_l2(out, ref)
    log(f"  image rel L2 vs reference {e:.2e}")
    assert e <= IMG_RL2


def assert_lik(lik, ref):
    """The reference sums the likelihood's terms (both signs) serially in float, which
    rounds by up to ~1e-4 relative on these 263k-bin sinograms; the device sums them
    in double.  NaN entries (est + additive < 0) must coincide."""
    lik = np.asarray(lik, np.float64)
    ref = np.asarray(ref, np.float64)
    assert np.array_equal(np.isnan(lik), np.isnan(ref)), (lik, ref)
    m = np.isfinite(ref)
    np.testing.assert_allclose(lik[m], ref[m], rtol=5e-4)


def mid_plan(ctx, g):
    return hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                           g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])


def mmr_plan(ctx, g):
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        297.0, 328.0, g["rings"], g["zfov"], g["span"], g["mrd"])
    p.set_scanner("mmr")
    return p


@pytest.mark.parametrize("mode", [hip.BP_GATHER, hip.BP_SCATTER])
def test_osem_with_additive_and_multiplicative_vs_reference(torch, ctx, golden, mode):
    f = golden("corrections.npz")
    g = json.loads(str(f["geom"]))
    p = mid_plan(ctx, g)
    meas_np = golden("mid3d_project_backproject.npz")["meas"].astype(np.float32)
    meas = dev(torch, meas_np.reshape(p.subset_shape()))
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    S, T = g["subsets"], g["iters"]
    for label, neg, use_mult in (("pos", False, True), ("neg", True, False)):
        log(f"case: mid3d OSEM {S}x{T} with corrections ({label}), mode {mode}")
        add, mult = phantoms.correction_sinograms(p.nsino, g["nproj"], g["nr"], negative=neg)
        sens = torch.empty((S,) + p.image_shape(), dtype=torch.float32, device="cuda")
        img, lik = p.reconstruct(meas, init, T, subsets=S, bp_mode=mode,
                                 additive=dev(torch, add),
                                 mult=dev(torch, mult) if use_mult else None, sens_out=sens)
        torch.cuda.synchronize()
        sh = host(sens)
        if use_mult:
            assert_voxel_close(sh[0], f[f"{label}_sens0"])
            assert_voxel_close(sh[3], f[f"{label}_sens3"])
        else:
            assert_voxel_close(sh[0], golden("mid3d_osem.npz")["sens0"])
        assert_image(host(img), f[f"{label}_final"])
        assert_lik(lik, f[f"{label}_lik"])
    # each bin's own additive angle (the documented alternative) is another image
    add, mult = phantoms.correction_sinograms(p.nsino, g["nproj"], g["nr"])
    own, _ = p.reconstruct(meas, init, T, subsets=S, bp_mode=mode, additive=dev(torch, add),
                           mult=dev(torch, mult), additive_own_angles=True)
    torch.cuda.synchronize()
    assert rel_l2(host(own), f["pos_final"]) > 1e-3


def test_fused_additive_indexing_and_no_clamp(torch, ctx, golden):
    """The fused forward step alone: subset bin (q, k, r) takes the additive bin
    (q, k, r) with APIRL_RATIO_ADDITIVE_LOCAL and (q, s + S k, r) without it; a sum
    below zero stays negative (ratio y / est < 0 where y != 0)."""
    log("case: fused projection + ratio, additive indexing")
    f = golden("corrections.npz")
    g = json.loads(str(f["geom"]))
    p = mid_plan(ctx, g)
    img = dev(torch, golden("mid3d_project_backproject.npz")["img"])
    meas_np = golden("mid3d_project_backproject.npz")["meas"].astype(np.float32)
    meas_np = meas_np.reshape(p.subset_shape())
    add, _ = phantoms.correction_sinograms(p.nsino, g["nproj"], g["nr"], negative=True)
    S, s = 7, 2
    est = host(p.project(img, subset=s, nsubsets=S)).astype(np.float32)
    for local in (True, False):
        r, _ = p.project_ratio(img, dev(torch, meas_np), additive=dev(torch, add), subset=s,
                               nsubsets=S, additive_local=local, want_lik=False)
        torch.cuda.synchronize()
        idx = np.arange(est.shape[1]) if local else s + S * np.arange(est.shape[1])
        tot = est + add[:, idx, :]
        y = meas_np[:, s::S, :]
        want = np.where((y != 0) & (tot != 0), y / np.where(tot != 0, tot, 1), 0).astype(np.float32)
        got = host(r)
        assert np.array_equal(got.view(np.int32), want.view(np.int32))
        assert (got < 0).any()  # est + additive < 0 is kept, not clamped


def test_mmr_subset_projection_bitexact(torch, ctx, golden):
    """Projection of subset 1 of 3 on the LORs of the reference's mMR subset
    sinograms (APIRL_SCANNER_MMR_SUBSET), bit-exact against the reference's."""
    log("case: mMR subset 1 of 3 projection")
    f = golden("corrections.npz")
    g = json.loads(str(f["geom_mmr"]))
    p = mmr_plan(ctx, g)
    mm = golden("mmr.npz")
    p.set_scanner("mmr_subset")
    out = p.project(dev(torch, mm["img"]), subset=1, nsubsets=3)
    torch.cuda.synchronize()
    got = host(out).ravel()
    assert np.array_equal(got.view(np.int32), f["mmr_subset1of3_proj"].view(np.int32))
    p.set_scanner("mmr")
    full_form = host(p.project(dev(torch, mm["img"]), subset=1, nsubsets=3)).ravel()
    assert not np.array_equal(full_form, got)


def mmr_osem_case(torch, ctx, golden):
    """The mMR OSEM case: a 192 mm image inside the radial FOV, seeded counts."""
    f = golden("corrections.npz")
    g = json.loads(str(f["geom_mmr_osem"]))
    p = mmr_plan(ctx, g)
    rng = np.random.default_rng(g["seed"])
    counts = rng.poisson(2.0, p.subset_shape()).astype(np.float32)
    assert hashlib.sha256(counts.tobytes()).hexdigest() == str(f["mmr_osem_meas_sha256"])
    return f, g, p, counts


@pytest.mark.parametrize("mode", [hip.BP_GATHER, hip.BP_SCATTER])
def test_mmr_osem_subsets_vs_reference(torch, ctx, golden, mode):
    """OSEM with 3 subsets on an mMR plan: the driver traces the subsets on the
    reference's subset LORs by itself and the final projection on the mMR's.  The
    likelihoods are not compared: the reference's projector reads before the pixel
    array for LORs that miss this image (SiddonProjector.cpp:417-423), so its values
    are undefined; the images are not affected (those LORs write nothing)."""
    log(f"case: mMR OSEM 3 subsets x 2, mode {mode}")
    f, g, p, counts = mmr_osem_case(torch, ctx, golden)
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    S = g["subsets"]
    sens = torch.empty((S,) + p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(dev(torch, counts), init, g["iters"], subsets=S, bp_mode=mode,
                             sens_out=sens)
    torch.cuda.synchronize()
    assert_voxel_close(host(sens)[0], f["mmr_osem_sens0"])
    out = host(img)
    assert_image(out, f["mmr_osem_final"])
    assert np.isfinite(out[:-1]).all()  # NaN only in the top slice, as the reference
    # the plan is back on the mMR LORs afterwards
    assert host(p.project(init, subset=1, nsubsets=3)).any()


def test_mmr_mlem3d_with_corrections_vs_reference(torch, ctx, golden):
    log("case: mMR MLEM3D with additive and multiplicative sinograms")
    f = golden("corrections.npz")
    g = json.loads(str(f["geom_mmr"]))
    p = mmr_plan(ctx, g)
    meas = dev(torch, golden("mmr.npz")["meas"].astype(np.float32).reshape(p.subset_shape()))
    add, mult = phantoms.correction_sinograms(p.nsino, g["nproj"], g["nr"])
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    sens = torch.empty(p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, g["mlem_iters"], subsets=1, bp_mode=hip.BP_GATHER,
                             osem_rules=0, additive=dev(torch, add), mult=dev(torch, mult),
                             sens_out=sens)
    torch.cuda.synchronize()
    assert_voxel_close(host(sens), f["mmr_mlem_sens"])
    assert_image(host(img), f["mmr_mlem_final"])
    np.testing.assert_allclose(np.asarray(lik, np.float64),
                               f["mmr_mlem_lik"].astype(np.float64), rtol=5e-4)


def test_batch_mmr_osem_and_corrections_equal_alone(torch, ctx, golden):
    """Batched jobs follow the same rules: an mMR OSEM job (subset LORs) and a
    mid-size OSEM job with both correction terms give, bit for bit, the images they
    give alone with the gather backprojector."""
    log("case: batch with mMR OSEM and corrected OSEM jobs")
    f, gm, pm, counts = mmr_osem_case(torch, ctx, golden)
    g = json.loads(str(f["geom"]))
    pc = mid_plan(ctx, g)
    mm_meas = dev(torch, counts)
    c_meas = dev(torch, golden("mid3d_project_backproject.npz")["meas"].astype(np.float32)
                 .reshape(pc.subset_shape()))
    add_np, mult_np = phantoms.correction_sinograms(pc.nsino, g["nproj"], g["nr"])
    add, mult = dev(torch, add_np), dev(torch, mult_np)
    ones_m = torch.ones(pm.image_shape(), dtype=torch.float32, device="cuda")
    ones_c = torch.ones(pc.image_shape(), dtype=torch.float32, device="cuda")
    alone_m, _ = pm.reconstruct(mm_meas, ones_m, 2, subsets=3, bp_mode=hip.Batch.backprojector(pm))
    assert np.isfinite(host(alone_m)).sum() > 0
    alone_c, _ = pc.reconstruct(c_meas, ones_c, 2, subsets=7, bp_mode=hip.Batch.backprojector(pc),
                                additive=add, mult=mult)
    b = hip.Batch(ctx, max_jobs=4, max_sinograms=1024)
    out_m = torch.empty_like(ones_m)
    out_c = torch.empty_like(ones_c)
    b.submit(pm, mm_meas, ones_m, out_m, 2, subsets=3)
    b.submit(pc, c_meas, ones_c, out_c, 2, subsets=7, additive=add, mult=mult)
    b.run()
    torch.cuda.synchronize()
    for a, o in ((alone_m, out_m), (alone_c, out_c)):
        x, y = host(a), host(o)
        assert np.array_equal(np.isnan(x), np.isnan(y))
        m = np.isfinite(x)
        assert np.array_equal(x[m].view(np.int32), y[m].view(np.int32))
    b.close()
