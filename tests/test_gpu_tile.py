"""The tile backprojector (APIRL_BP_TILE) through the C-ABI.

A wavefront owns 8 x 8 voxel columns by a slab of slices and walks, one sinogram per
lane, the reference's Siddon loop of every ray that crosses them from the tile's
entry to its exit; its terms are the reference's floats, summed in double in LDS.

Bars: every voxel within 1e-5 relative (floor 1e-6 of the image maximum) of the
reference's or the oracle's backprojection, as for the other forms; bit-identical
from run to run; the backprojection of a single bin (one term per voxel: the
reference's float weight times the bin) bit-exact against the gather, whose terms
are pinned to the oracle's; the fused EM update bit-identical to the backprojection
followed by the update; images after N iterations within 1e-4 relative L2."""
import json

import numpy as np
import pytest

import _oracle as O
from apirl_amd import hip, phantoms
from test_gpu_parity import (assert_bitexact, assert_voxel_close, ctx, dev, host, log, plan3d,  # noqa: F401
                             rel_l2, torch)

pytestmark = pytest.mark.gpu

C2 = dict(nx=256, ny=256, nz=81, dx=2.0445, dz=2.027, nproj=252, nr=344, rings=64,
          rfov=261.696, zfov=164.187, rscanner=400.0, span=11, mrd=60)


def test_tile_mid3d_vs_reference_deterministic(torch, ctx, golden):
    log("case: mid3d tile backprojection, slab depths 0 (default), 4, 17")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32)).reshape(p.subset_shape())
    for zs in (0, 4, 17):
        if zs:
            p.set_tuning(tile_zs=zs)
        out = p.backproject(meas, mode=hip.BP_TILE)
        out2 = p.backproject(meas, mode=hip.BP_TILE)
        torch.cuda.synchronize()
        assert_bitexact(host(out2), host(out))
        assert_voxel_close(host(out), f["back"])
    assert p.get_tuning()["tile_zs"] == 17


def test_tile_single_bins_bitexact_vs_gather(torch, ctx):
    """One non-zero bin at a time (seeded, plus the subset's first and last bins) on
    the 256x256x81 geometry, subset 3 of 21: each voxel receives one term, so the
    tile form and the gather agree bit for bit on every voxel."""
    log("case: c2 single-bin tile backprojections vs gather")
    p = plan3d(ctx, C2)
    S, s = 21, 3
    shape = p.subset_shape(s, S)
    rng = np.random.default_rng(11)
    picks = [np.unravel_index(int(q), shape) for q in rng.choice(int(np.prod(shape)), 14, replace=False)]
    picks += [(0, 0, shape[2] // 2), (shape[0] - 1, shape[1] - 1, shape[2] // 2)]
    for b in picks:
        sino = np.zeros(shape, np.float32)
        sino[b] = rng.uniform(0.5, 3.0)
        st = dev(torch, sino)
        t = p.backproject(st, subset=s, nsubsets=S, mode=hip.BP_TILE)
        gth = p.backproject(st, subset=s, nsubsets=S, mode=hip.BP_GATHER)
        torch.cuda.synchronize()
        assert_bitexact(host(t), host(gth))


def test_tile_c2_subset_vs_oracle(torch, ctx, golden):
    log("case: c2 tile backprojection of subset 0/21 vs oracle")
    f = golden("c2_rays.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    img, _ = phantoms.cylinder_points3d(g["nx"], g["nz"], g["dx"], g["dz"])
    sino = p.project(dev(torch, img), subset=0, nsubsets=21)
    out = p.backproject(sino, subset=0, nsubsets=21, mode=hip.BP_TILE)
    out2 = p.backproject(sino, subset=0, nsubsets=21, mode=hip.BP_TILE)
    torch.cuda.synchronize()
    assert_bitexact(host(out2), host(out))
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    ang, r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    idx = O.subset_angles(g["nproj"], 21, 0)
    ref = O.backproject3d(im, host(sino), ang[idx], r, g["rscanner"], za, zb)
    assert_voxel_close(host(out), ref)


def test_tile_divide_and_sensitivity_vs_reference(torch, ctx, golden):
    """DivideAndBackproject (every zero combination of measured and estimated bins)
    and the subset sensitivities against the reference; the multiplicative-sinogram
    sensitivity against the gather's."""
    log("case: mid3d tile divide-and-backproject and sensitivity")
    d = golden("divide_backproject.npz")
    g = json.loads(str(d["geom3d"]))
    p = plan3d(ctx, g)
    meas = dev(torch, d["meas3d"].astype(np.float32).reshape(p.subset_shape()))
    est = dev(torch, (d["est3d_x2"].astype(np.float32) * np.float32(0.5)).reshape(p.subset_shape()))
    div = p.divide_backproject(meas, est, mode=hip.BP_TILE)
    torch.cuda.synchronize()
    assert_voxel_close(host(div).ravel(), d["img3d"])
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    S = g["subsets"]
    s0 = p.sensitivity(subset=0, nsubsets=S, mode=hip.BP_TILE)
    s6 = p.sensitivity(subset=6, nsubsets=S, mode=hip.BP_TILE)
    mult = dev(torch, np.random.default_rng(3).uniform(0.5, 1.5, p.subset_shape()).astype(np.float32))
    m_t = p.sensitivity(mult=mult, subset=2, nsubsets=S, mode=hip.BP_TILE)
    m_g = p.sensitivity(mult=mult, subset=2, nsubsets=S, mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_voxel_close(host(s0), f["sens0"])
    assert_voxel_close(host(s6), f["sens6"])
    assert_voxel_close(host(m_t), host(m_g))


@pytest.mark.parametrize("strict", [0, 1])
def test_tile_fused_update_equals_backproject_then_update(torch, ctx, strict):
    log(f"case: tile backprojection with the EM update fused in, strict={strict}")
    g = C2
    p = plan3d(ctx, g)
    rng = np.random.default_rng(5)
    S, s = 21, 8
    img = phantoms.cylinder3d(g["nx"], g["nz"], g["dx"], g["dz"], radius=0.6 * g["nx"] * g["dx"] / 2)
    sino = p.project(dev(torch, img), subset=s, nsubsets=S)
    sens = p.backproject(torch.ones_like(sino), subset=s, nsubsets=S, mode=hip.BP_TILE)
    thr = hip.mlem_threshold(ctx, sens)
    x0 = dev(torch, rng.uniform(0.5, 2.0, img.shape))
    b = torch.zeros_like(x0)
    p.backproject(sino, out=b, subset=s, nsubsets=S, mode=hip.BP_TILE)
    ref = hip.em_update(ctx, x0.clone(), b, sens, thr, strict)
    got = p.backproject_update(sino, x0.clone(), sens, thr, strict, subset=s, nsubsets=S,
                               mode=hip.BP_TILE)
    torch.cuda.synchronize()
    r, o = host(ref), host(got)
    assert (o == 0).sum() > 0 and (o > 0).sum() > 0
    assert np.array_equal(o.view(np.int32), r.view(np.int32))


def test_tile_kept_geometry_equals_rebuilt(torch, ctx, golden):
    """The geometry a plan keeps per subset (apirl_tuning.tile_cache_mb) gives the same
    images, bit for bit, as geometry rebuilt on every call: subsets called in turn, twice,
    with sinograms whose zero bins (whole rays and scattered bins) differ from call to
    call, a multiplicative sensitivity and the fused update; each within the
    backprojection bar of the gather."""
    log("case: mid3d tile backprojection with kept vs rebuilt geometry")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    kept, rebuilt = plan3d(ctx, g), plan3d(ctx, g)
    rebuilt.set_tuning(tile_cache_mb=-1)
    assert rebuilt.get_tuning()["tile_cache_mb"] == 0
    assert kept.get_tuning()["tile_cache_mb"] == 16384
    S = 3
    rng = np.random.default_rng(21)
    for rep in range(2):
        for s in range(S):
            shape = kept.subset_shape(s, S)
            sino = rng.uniform(0.5, 2.0, shape).astype(np.float32)
            sino[:, :, rng.random(shape[2]) < 0.3] = 0.0
            sino[rng.random(shape) < 0.2] = 0.0
            st = dev(torch, sino)
            a = kept.backproject(st, subset=s, nsubsets=S, mode=hip.BP_TILE)
            b = rebuilt.backproject(st, subset=s, nsubsets=S, mode=hip.BP_TILE)
            gth = kept.backproject(st, subset=s, nsubsets=S, mode=hip.BP_GATHER)
            torch.cuda.synchronize()
            assert_bitexact(host(a), host(b))
            assert_voxel_close(host(a), host(gth))
    mult = dev(torch, rng.uniform(0.5, 1.5, kept.subset_shape()).astype(np.float32))
    for s in range(S):
        a = kept.sensitivity(mult=mult, subset=s, nsubsets=S, mode=hip.BP_TILE)
        b = rebuilt.sensitivity(mult=mult, subset=s, nsubsets=S, mode=hip.BP_TILE)
        torch.cuda.synchronize()
        assert_bitexact(host(a), host(b))
    s = 1
    sino = dev(torch, rng.uniform(0.0, 2.0, kept.subset_shape(s, S)).astype(np.float32))
    sens = kept.sensitivity(subset=s, nsubsets=S, mode=hip.BP_TILE)
    thr = hip.mlem_threshold(ctx, sens)
    x0 = dev(torch, rng.uniform(0.5, 2.0, kept.image_shape()).astype(np.float32))
    a = kept.backproject_update(sino, x0.clone(), sens, thr, 0, subset=s, nsubsets=S,
                                mode=hip.BP_TILE)
    b = rebuilt.backproject_update(sino, x0.clone(), sens, thr, 0, subset=s, nsubsets=S,
                                   mode=hip.BP_TILE)
    torch.cuda.synchronize()
    assert_bitexact(host(a), host(b))


def test_tile_osem_mid3d_vs_reference(torch, ctx, golden):
    log("case: mid3d OSEM 7x2 with the tile backprojector")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32))
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, g["iters"], subsets=g["subsets"], bp_mode=hip.BP_TILE)
    img2, _ = p.reconstruct(meas, init, g["iters"], subsets=g["subsets"], bp_mode=hip.BP_TILE)
    torch.cuda.synchronize()
    assert_bitexact(host(img2), host(img), nan_ok=True)
    out, ref = host(img), f["final"]
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    assert rel_l2(out, ref) <= 1e-4


def test_tile_mmr_and_oversampled_vs_reference(torch, ctx, golden):
    """The mMR geometry (non-uniform radial grid, odd bins at half-step angles: the
    candidate search through its lookup table) and the (3, 2) oversampled LORs (the
    six LORs of a bin backprojected in turn)."""
    log("case: tile backprojection on the mMR geometry and with oversampled LORs")
    f = golden("mmr.npz")
    g = json.loads(str(f["geom"]))
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        297.0, 328.0, g["rings"], g["zfov"], g["span"], g["mrd"])
    p.set_scanner("mmr")
    meas = dev(torch, f["meas"].astype(np.float32).reshape(p.subset_shape()))
    t1 = p.backproject(meas, mode=hip.BP_TILE)
    t2 = p.backproject(meas, mode=hip.BP_TILE)
    torch.cuda.synchronize()
    assert_voxel_close(host(t1), f["back"])
    assert_bitexact(host(t2), host(t1))
    o = golden("oversampled.npz")
    g = json.loads(str(o["geom"]))
    m = golden("mid3d_project_backproject.npz")
    p = plan3d(ctx, g)
    p.set_oversampling(g["ns"], g["nas"])
    meas = dev(torch, m["meas"].astype(np.float32).reshape(p.subset_shape()))
    back = p.backproject(meas, mode=hip.BP_TILE)
    torch.cuda.synchronize()
    assert_voxel_close(host(back), o["back"])


def test_tile_refusals(torch, ctx, golden):
    log("case: tile backprojector refusals")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32)).reshape(p.subset_shape())
    with pytest.raises(hip.ApirlError, match="reference precision"):
        p.backproject(meas, mode=hip.BP_TILE, precision=hip.PREC_KAHAN)
    c = golden("c1_project_backproject.npz")
    g2 = json.loads(str(c["geom"]))
    from test_gpu_parity import c1_plan
    p2 = c1_plan(ctx, g2)
    s2 = torch.zeros(p2.subset_shape(), dtype=torch.float32, device="cuda")
    with pytest.raises(hip.ApirlError, match="3D only"):
        p2.backproject(s2, mode=hip.BP_TILE)
