"""Parity of the HIP path (through the C-ABI) with the oracle and the reference's
goldens.

Bars: forward projections and every integer / index quantity bit-exact; the scatter
backprojector (fp32 atomics, order not fixed) within 1e-5 relative per voxel; the
gather backprojector bit-exact in 2D, within 1e-5 relative per voxel in 3D (fixed
order, bit-identical from run to run); reconstructed images within 1e-4 relative L2."""
import json
import sys

import numpy as np
import pytest

import _oracle as O
from apirl_amd import hip, phantoms

pytestmark = pytest.mark.gpu

# per-voxel relative tolerance of the scatter backprojector, with a floor at 1e-6 of
# the image maximum for voxels whose value is a tiny edge sum
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


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def assert_bitexact(a, b, nan_ok=False):
    a = np.ascontiguousarray(a, np.float32).ravel()
    b = np.ascontiguousarray(b, np.float32).ravel()
    assert a.shape == b.shape
    if nan_ok:
        na, nb = np.isnan(a), np.isnan(b)
        assert np.array_equal(na, nb)
        a, b = a[~na], b[~nb]
    d = np.flatnonzero(bits(a) != bits(b))
    assert d.size == 0, f"{d.size} of {a.size} differ, first {d[:5]}: {a[d[:5]]} vs {b[d[:5]]}"


def assert_voxel_close(a, b, rtol=BP_RTOL, floor=BP_FLOOR):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    scale = np.maximum(np.abs(b), floor * np.abs(b).max())
    err = np.abs(a - b) / scale
    worst = int(np.argmax(err))
    assert err.max() <= rtol, f"max rel err {err.max():.3e} at {worst}: {a[worst]} vs {b[worst]}"


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    m = np.isfinite(a) & np.isfinite(b)
    return np.linalg.norm(a[m] - b[m]) / np.linalg.norm(b[m])


def c1_plan(ctx, g):
    return hip.Plan.for_2d(ctx, g["nx"], g["dx"], g["nproj"], g["nr"], g["rfov"], g["rscanner"])


def plan3d(ctx, g):
    return hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                           g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])


# ------------------------------------------------------------------------ 2D

def test_c1_project_bitexact_vs_reference(torch, ctx, golden):
    log("case: c1 2D projection 128x128 / 192x192")
    f = golden("c1_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = c1_plan(ctx, g)
    out = p.project(dev(torch, f["img"]).reshape(1, 128, 128))
    torch.cuda.synchronize()
    assert_bitexact(host(out), f["proj"])


def test_c1_backproject_scatter_vs_reference(torch, ctx, golden):
    log("case: c1 2D scatter backprojection")
    f = golden("c1_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = c1_plan(ctx, g)
    out = p.backproject(dev(torch, f["meas"].astype(np.float32)).reshape(1, 192, 192))
    torch.cuda.synchronize()
    assert_voxel_close(host(out), f["back"])


def test_c1_mlem2d_60_iterations_vs_reference(torch, ctx, golden):
    log("case: c1 MLEM 2D 60 iterations")
    f = golden("c1_mlem2d.npz")
    g = json.loads(str(f["geom"]))
    p = c1_plan(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32)).reshape(1, 192, 192)
    init = torch.ones((1, 128, 128), dtype=torch.float32, device="cuda")
    sens = torch.empty((1, 128, 128), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, g["iters"], subsets=1, sens_out=sens)
    torch.cuda.synchronize()
    assert_voxel_close(host(sens), f["sens"])
    e = rel_l2(host(img), f["final"])
    log(f"c1 MLEM rel L2 vs reference {e:.2e}")
    assert e <= IMG_RL2
    ref_lik = f["lik"].astype(np.float64)
    assert np.allclose(lik[:-1], ref_lik[:-1], rtol=1e-4), (lik[:5], ref_lik[:5])


# ------------------------------------------------------------------------ 3D

def test_mid3d_project_bitexact_vs_reference(torch, ctx, golden):
    log("case: mid3d projection 64x64x17 / 86x42x81")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    out = p.project(dev(torch, f["img"]))
    torch.cuda.synchronize()
    assert_bitexact(host(out), f["proj"])


def test_mid3d_project_subsets_bitexact_vs_oracle(torch, ctx, golden):
    log("case: mid3d subset projections")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    full = f["proj"].reshape(p.nsino, g["nproj"], g["nr"])
    img = dev(torch, f["img"])
    for S in (7, 5):
        for s in range(S):
            out = p.project(img, subset=s, nsubsets=S)
            idx = O.subset_angles(g["nproj"], S, s)
            assert_bitexact(host(out), full[:, idx, :])


def test_mid3d_backproject_scatter_vs_reference(torch, ctx, golden):
    log("case: mid3d scatter backprojection")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32))
    out = p.backproject(meas.reshape(p.subset_shape()))
    torch.cuda.synchronize()
    assert_voxel_close(host(out), f["back"])


def test_mid3d_divide_backproject_vs_oracle(torch, ctx, golden):
    log("case: mid3d divide-and-backproject")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    ang, r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    meas = f["meas"].astype(np.float32).reshape(len(za), g["nproj"], g["nr"])
    est = f["proj"].reshape(meas.shape).copy()
    est[:, :, ::7] = 0  # exercise the est == 0 branch
    ref = O.divide_backproject3d(im, meas, est, ang, r, g["rscanner"], za, zb)
    out = p.divide_backproject(dev(torch, meas), dev(torch, est))
    torch.cuda.synchronize()
    assert_voxel_close(host(out), ref)


@pytest.mark.parametrize("dims", [2, 3])
def test_divide_backproject_vs_reference(torch, ctx, golden, dims):
    """DivideAndBackproject against the reference's own output, with measured and
    estimated bins zero in every combination (2D: C1 geometry, 3D: mid-size), in
    both forms; the gather form is bit-identical from run to run."""
    log(f"case: {dims}D divide-and-backproject vs reference")
    f = golden("divide_backproject.npz")
    g = json.loads(str(f[f"geom{dims}d"]))
    p = c1_plan(ctx, g) if dims == 2 else plan3d(ctx, g)
    meas = f[f"meas{dims}d"].astype(np.float32)
    est = f[f"est{dims}d_x2"].astype(np.float32) * np.float32(0.5)
    shape = (1,) + meas.shape if dims == 2 else meas.shape
    m, e = dev(torch, meas.reshape(shape)), dev(torch, est.reshape(shape))
    out = p.divide_backproject(m, e)
    gat = p.divide_backproject(m, e, mode=hip.BP_GATHER)
    gat2 = p.divide_backproject(m, e, mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_voxel_close(host(out).ravel(), f[f"img{dims}d"])
    assert_voxel_close(host(gat).ravel(), f[f"img{dims}d"])
    assert_bitexact(host(gat2), host(gat))


def test_mid3d_fused_ratio_and_likelihood(torch, ctx, golden):
    log("case: mid3d fused projection + ratio + likelihood")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    full = f["proj"].reshape(p.nsino, g["nproj"], g["nr"])
    meas = f["meas"].astype(np.float32).reshape(full.shape)
    add = np.full(full.shape, 0.25, np.float32)
    img = dev(torch, f["img"])
    S, s = 7, 3
    idx = O.subset_angles(g["nproj"], S, s)
    ratio, lik = p.project_ratio(img, dev(torch, meas), dev(torch, add), subset=s, nsubsets=S)
    torch.cuda.synchronize()
    est = (full[:, idx, :] + add[:, idx, :]).astype(np.float32)
    y = meas[:, idx, :]
    assert_bitexact(host(ratio), O.inverse_divide(est, y).reshape(est.shape))
    m = est != 0
    ref = np.sum(y[m].astype(np.float64) * np.log(est[m].astype(np.float64)) - est[m])
    assert abs(float(host(lik)[0]) - ref) <= 1e-6 * abs(ref) + 1e-3


def test_mid3d_sensitivity_and_update(torch, ctx, golden):
    log("case: mid3d sensitivity, threshold, EM update")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    sens0 = p.sensitivity(subset=0, nsubsets=g["subsets"])
    torch.cuda.synchronize()
    assert_voxel_close(host(sens0), f["sens0"])
    s_np = f["sens0"]
    thr = hip.mlem_threshold(ctx, dev(torch, s_np))
    assert np.float32(thr) == np.float32(O.mlem_threshold(s_np))
    rng = np.random.default_rng(3)
    x = rng.random(s_np.shape, dtype=np.float32)
    b = rng.random(s_np.shape, dtype=np.float32)
    for strict in (0, 1):
        xd = dev(torch, x)
        hip.em_update(ctx, xd, dev(torch, b), dev(torch, s_np), thr, strict)
        torch.cuda.synchronize()
        assert_bitexact(host(xd), O.em_update(x, b, s_np, thr, strict), nan_ok=True)


def test_mid3d_osem_vs_reference(torch, ctx, golden):
    log("case: mid3d OSEM 7 subsets x 2 iterations")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32))
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, g["iters"], subsets=g["subsets"])
    torch.cuda.synchronize()
    out = host(img)
    ref = f["final"]
    assert np.array_equal(np.isnan(out), np.isnan(ref)), "NaN positions differ"
    e = rel_l2(out, ref)
    log(f"mid3d OSEM rel L2 vs reference {e:.2e}")
    assert e <= IMG_RL2
    assert np.allclose(lik[:-1], f["lik"][:-1].astype(np.float64), rtol=1e-4)


def test_c2_full_projection_sampled_bins_bitexact(torch, ctx, golden):
    """256x256x81 image, 344 r x 252 angles x 837 sinograms: the full projection on
    the GPU, checked bit-exact at the reference's sampled rays."""
    log("case: c2 full projection 256x256x81 / 344x252x837")
    f = golden("c2_rays.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    img, _ = phantoms.cylinder_points3d(g["nx"], g["nz"], g["dx"], g["dz"])
    out = p.project(dev(torch, img))
    torch.cuda.synchronize()
    flat = out.reshape(-1)
    bins = torch.arange(g["start"], flat.numel(), g["stride"], device="cuda")
    assert_bitexact(host(flat[bins]), f["proj"])
    n_int = p.intersections()
    log(f"c2 intersections per full projection: {n_int}")
    assert n_int > 1e10
    # checksum of checksums across subsets: 21 subset projections tile the full one
    tot = 0.0
    for s in range(21):
        tot += float(p.project(dev(torch, img), subset=s, nsubsets=21).double().sum())
    assert abs(tot - float(flat.double().sum())) <= 1e-9 * abs(tot)


# ------------------------------------------------------------ gather (deterministic)

def test_c1_backproject_gather_bitexact_vs_reference(torch, ctx, golden):
    log("case: c1 2D gather backprojection")
    f = golden("c1_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = c1_plan(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32)).reshape(1, 192, 192)
    out = p.backproject(meas, mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_bitexact(host(out), f["back"])


def test_mid3d_backproject_gather_deterministic_and_close_to_reference(torch, ctx, golden):
    log("case: mid3d gather backprojection")
    f = golden("mid3d_project_backproject.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32)).reshape(p.subset_shape())
    out = p.backproject(meas, mode=hip.BP_GATHER)
    out2 = p.backproject(meas, mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_bitexact(host(out2), host(out))  # deterministic
    assert_voxel_close(host(out), f["back"])


def test_mid3d_sensitivity_gather_close_to_reference(torch, ctx, golden):
    log("case: mid3d gather sensitivity")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    s0 = p.sensitivity(subset=0, nsubsets=g["subsets"], mode=hip.BP_GATHER)
    s6 = p.sensitivity(subset=6, nsubsets=g["subsets"], mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_voxel_close(host(s0), f["sens0"])
    assert_voxel_close(host(s6), f["sens6"])


def test_c1_mlem2d_gather_bitexact_vs_reference(torch, ctx, golden):
    log("case: c1 MLEM 2D 60 iterations, gather")
    f = golden("c1_mlem2d.npz")
    g = json.loads(str(f["geom"]))
    p = c1_plan(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32)).reshape(1, 192, 192)
    init = torch.ones((1, 128, 128), dtype=torch.float32, device="cuda")
    sens = torch.empty((1, 128, 128), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, g["iters"], subsets=1, bp_mode=hip.BP_GATHER,
                             sens_out=sens)
    torch.cuda.synchronize()
    assert_bitexact(host(sens), f["sens"])
    assert_bitexact(host(img), f["final"])


def test_mid3d_osem_gather_deterministic_and_close_to_reference(torch, ctx, golden):
    log("case: mid3d OSEM gather 7x2")
    f = golden("mid3d_osem.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    meas = dev(torch, f["meas"].astype(np.float32))
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, g["iters"], subsets=g["subsets"], bp_mode=hip.BP_GATHER)
    img2, _ = p.reconstruct(meas, init, g["iters"], subsets=g["subsets"], bp_mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_bitexact(host(img2), host(img), nan_ok=True)  # bit-for-bit deterministic
    out, ref = host(img), f["final"]
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    assert rel_l2(out, ref) <= IMG_RL2


def test_oversampled_projector_vs_reference(torch, ctx, golden):
    """Three radial x two axial LORs per bin (apirl_plan_set_oversampling): the
    projection is bit-exact against the reference's SiddonProjector(3, 2), the scatter
    and gather backprojections and DivideAndBackproject within the per-voxel bar (the
    gather run-to-run bit-identical: the six LORs of a bin gathered in turn); one LOR
    per bin restores the plain projector bit for bit."""
    log("case: oversampled projector 3 x 2, mid3d")
    f = golden("oversampled.npz")
    g = json.loads(str(f["geom"]))
    m = golden("mid3d_project_backproject.npz")
    p = plan3d(ctx, g)
    p.set_oversampling(g["ns"], g["nas"])
    proj = p.project(dev(torch, m["img"]))
    meas = dev(torch, m["meas"].astype(np.float32).reshape(p.subset_shape()))
    back = p.backproject(meas, mode=hip.BP_SCATTER)
    d = golden("divide_backproject.npz")
    dmeas = dev(torch, d["meas3d"].astype(np.float32).reshape(p.subset_shape()))
    dest = dev(torch, (d["est3d_x2"].astype(np.float32) * np.float32(0.5))
               .reshape(p.subset_shape()))
    div = p.divide_backproject(dmeas, dest, mode=hip.BP_SCATTER)
    torch.cuda.synchronize()
    assert_bitexact(host(proj).ravel(), f["proj"])
    assert_voxel_close(host(back), f["back"])
    assert_voxel_close(host(div), f["divback"])
    back_g = host(p.backproject(meas, mode=hip.BP_GATHER))
    assert np.array_equal(back_g.view(np.int32),
                          host(p.backproject(meas, mode=hip.BP_GATHER)).view(np.int32))
    assert_voxel_close(back_g, f["back"])
    assert_voxel_close(host(p.divide_backproject(dmeas, dest, mode=hip.BP_GATHER)),
                       f["divback"])
    with pytest.raises(hip.ApirlError, match="reference precision"):
        p.backproject(meas, mode=hip.BP_GATHER, precision=hip.PREC_KAHAN)
    with pytest.raises(hip.ApirlError, match="reference precision"):
        p.project(dev(torch, m["img"]), precision=hip.PREC_KAHAN)
    # OSEM with the oversampled projector runs end to end with either backprojector;
    # the two differ only in the order of the float additions per voxel
    init = torch.ones(p.image_shape(), dtype=torch.float32, device="cuda")
    img, lik = p.reconstruct(meas, init, 1, subsets=7, bp_mode=hip.BP_SCATTER)
    img_g, lik_g = p.reconstruct(meas, init, 1, subsets=7, bp_mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    out, out_g = host(img), host(img_g)
    assert np.isfinite(lik).all() and np.nanmax(out) > 0
    assert np.array_equal(np.isnan(out), np.isnan(out_g))
    fin = np.isfinite(out)
    assert np.linalg.norm(out_g[fin] - out[fin]) <= 1e-5 * np.linalg.norm(out[fin])
    p.set_oversampling(1, 1)
    plain = p.project(dev(torch, m["img"]))
    torch.cuda.synchronize()
    assert_bitexact(host(plain).ravel(), m["proj"])


def test_mmr_geometry_vs_reference(torch, ctx, golden):
    """apirl_plan_set_scanner(APIRL_SCANNER_MMR): projection bit-exact, scatter and
    gather backprojections within the per-voxel bar against the reference's mMR
    sinogram; the cylindrical geometry (and its gather) comes back unchanged."""
    log("case: mMR geometry, 24 x 86 x 251 sinograms")
    f = golden("mmr.npz")
    g = json.loads(str(f["geom"]))
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        297.0, 328.0, g["rings"], g["zfov"], g["span"], g["mrd"])
    ref_plain = p.project(dev(torch, f["img"]))
    p.set_scanner("mmr")
    proj = p.project(dev(torch, f["img"]))
    meas = dev(torch, f["meas"].astype(np.float32).reshape(p.subset_shape()))
    back = p.backproject(meas, mode=hip.BP_SCATTER)
    # the gather on the mMR's non-uniform grid and half-step odd angles: every voxel
    # within the per-voxel bar, and the same bits on a second run
    gat = p.backproject(meas, mode=hip.BP_GATHER)
    gat2 = p.backproject(meas, mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_bitexact(host(proj).ravel(), f["proj"])
    assert_voxel_close(host(back), f["back"])
    assert_voxel_close(host(gat), f["back"])
    assert_bitexact(host(gat2), host(gat))
    p.set_scanner("cylindrical")
    again = p.project(dev(torch, f["img"]))
    plain_gat = p.backproject(again, mode=hip.BP_GATHER)
    p.set_scanner("mmr")
    p.set_scanner("cylindrical")
    plain_gat2 = p.backproject(again, mode=hip.BP_GATHER)
    torch.cuda.synchronize()
    assert_bitexact(host(again), host(ref_plain))
    assert_bitexact(host(plain_gat2), host(plain_gat))


def test_mlem3d_on_mmr_vs_reference(torch, ctx, golden):
    """The device-resident driver's 3D MLEM (one subset, MlemSinogram3d's rules) on
    the mMR geometry against the reference's MlemSinogram3d run: sensitivity within
    the per-voxel bar, the image after 3 iterations within the image bar, the
    likelihood trace as stated below."""
    log("case: MLEM3D on mMR, 3 iterations")
    f = golden("mmr_mlem3d.npz")
    g = json.loads(str(f["geom"]))
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        297.0, 328.0, g["rings"], g["zfov"], g["span"], g["mrd"])
    p.set_scanner("mmr")
    meas = golden("mmr.npz")["meas"].astype(np.float32).reshape(p.subset_shape())
    init = np.ones((g["nz"], g["ny"], g["nx"]), np.float32)
    sens = p.sensitivity(mode=hip.BP_SCATTER)
    img, lik = p.reconstruct(dev(torch, meas), dev(torch, init), g["iters"], subsets=1,
                             bp_mode=hip.BP_SCATTER, osem_rules=0)
    torch.cuda.synchronize()
    assert_voxel_close(host(sens), f["sens"])
    got = host(img)
    assert np.isfinite(got).all()
    assert rel_l2(got, f["final"]) <= IMG_RL2
    # Likelihood: the reference adds its 518,064 float terms (of both signs) serially
    # in float, which here rounds by up to 1.3e-4 relative; the device sums them in
    # double.  Against the same terms summed in double (the oracle's projections of
    # the initial and of the reference's final image) the bar is 1e-5, against the
    # reference's float sums 5e-4.
    lik = np.asarray(lik, np.float64)
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    pts = O.mmr_lor_table(g["nproj"], g["nr"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    for t, x in ((0, init), (g["iters"], f["final"])):
        est = O.project3d_pts(im, x, pts, za, zb).astype(np.float64).ravel()
        y = meas.astype(np.float64).ravel()
        nz = est != 0
        exact = np.sum(y[nz] * np.log(est[nz]) - est[nz])
        assert abs(lik[t] - exact) <= 1e-5 * abs(exact), (t, lik[t], exact)
    np.testing.assert_allclose(lik, f["lik"].astype(np.float64), rtol=5e-4)
    # the same reconstruction with the gather backprojector
    sens_g = p.sensitivity(mode=hip.BP_GATHER)
    img_g, _ = p.reconstruct(dev(torch, meas), dev(torch, init), g["iters"], subsets=1,
                             bp_mode=hip.BP_GATHER, osem_rules=0)
    torch.cuda.synchronize()
    assert_voxel_close(host(sens_g), f["sens"])
    assert rel_l2(host(img_g), f["final"]) <= IMG_RL2


@pytest.mark.parametrize("which", ["mid3d", "c2"])
def test_gather_voxel_mappings_bit_identical(torch, ctx, golden, which):
    """The 3D gather's launch shapes give every voxel the same terms in the same
    order: the per-voxel kernel (gather_zlanes = 2) and the column kernel with 1-4
    voxels per lane (gather_zvox), in grid order or in XCD sectors (gather_xcd), agree
    bit for bit, in both float precision modes, on a sinogram with zero bins."""
    log(f"case: gather voxel mappings, {which}")
    if which == "mid3d":
        g = json.loads(str(golden("mid3d_project_backproject.npz")["geom"]))
        nsub = 7
    else:
        g = json.loads(str(golden("c2_rays.npz")["geom"]))
        nsub = 21
    p = plan3d(ctx, g)
    rng = np.random.default_rng(17)
    sino = rng.random(p.subset_shape(1, nsub), dtype=np.float32)
    sino[rng.random(sino.shape) < 0.3] = 0.0
    sd = dev(torch, sino)
    for prec in (hip.PREC_REFERENCE, hip.PREC_KAHAN):
        p.gather_zlanes = 2
        ref = host(p.backproject(sd, subset=1, nsubsets=nsub, mode=hip.BP_GATHER, precision=prec))
        p.gather_zlanes = 1
        for xcd in (hip.GATHER_XCD_ROUND_ROBIN, hip.GATHER_XCD_WEDGES):
            p.set_tuning(gather_xcd=xcd)
            for zv in (1, 2, 3, 4):
                p.gather_zvox = zv
                out = p.backproject(sd, subset=1, nsubsets=nsub, mode=hip.BP_GATHER,
                                    precision=prec)
                torch.cuda.synchronize()
                assert_bitexact(host(out), ref)
    p.set_tuning(gather_zvox=3, gather_xcd=hip.GATHER_XCD_ROUND_ROBIN)


def test_c2_subset_gather_vs_oracle(torch, ctx, golden):
    """256x256x81 / 837 sinograms, subset 0 of 21: gather and scatter backprojection
    of the phantom's projection, each voxel within 1e-5 relative of the oracle."""
    log("case: c2 gather backprojection of subset 0/21 vs oracle")
    f = golden("c2_rays.npz")
    g = json.loads(str(f["geom"]))
    p = plan3d(ctx, g)
    img, _ = phantoms.cylinder_points3d(g["nx"], g["nz"], g["dx"], g["dz"])
    sino = p.project(dev(torch, img), subset=0, nsubsets=21)
    out = p.backproject(sino, subset=0, nsubsets=21, mode=hip.BP_GATHER)
    sc = p.backproject(sino, subset=0, nsubsets=21, mode=hip.BP_SCATTER)
    torch.cuda.synchronize()
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    ang, r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    idx = O.subset_angles(g["nproj"], 21, 0)
    ref = O.backproject3d(im, host(sino), ang[idx], r, g["rscanner"], za, zb)
    assert_voxel_close(host(out), ref)
    assert_voxel_close(host(sc), ref)


def test_gather_refuses_subsets_beyond_32bit_offsets(torch, ctx):
    """The gather forms record offsets in 32 bits: a subset of 2^31 bins or more
    (2048 sinograms x 1024 angles x 1024 bins, one subset) is refused with an error
    before any scratch is allocated, by the single-job path and by batches."""
    log("case: gather refusal above 2^31 subset bins")
    nsino = 2048
    z = np.full(nsino, 1.0, np.float32)
    p = hip.Plan(ctx, 8, 8, 2, 4.0, 1.0, 1024, 1024, 16.0, 400.0, dims=3, z1=z, z2=z,
                 ring_width=1.0)
    img = torch.zeros(p.image_shape(), dtype=torch.float32, device="cuda")
    sino = torch.zeros(p.subset_shape(), dtype=torch.float32, device="cuda")  # 8 GiB
    with pytest.raises(hip.ApirlError, match="2\\^31"):
        p.backproject(sino, out=img, mode=hip.BP_GATHER)
    b = hip.Batch(ctx, max_jobs=2, max_sinograms=4096)
    with pytest.raises(hip.ApirlError, match="2\\^31"):
        b.submit(p, sino, img, img, 1, 1)
    b.close()
    del sino


@pytest.mark.parametrize("geom,strict", [("mid", 0), ("mid", 1), ("c2", 0)])
def test_backproject_update_equals_backproject_then_update(torch, ctx, geom, strict):
    """apirl_backproject_update (the gather with the EM update fused in, as the EM
    driver and the bench run it) equals the gather backprojection into a zeroed image
    followed by apirl_em_update, bit for bit, with either threshold rule and voxels
    below the threshold."""
    g = {"mid": dict(nx=64, ny=64, nz=17, dx=2.0445, dz=2.027, nproj=84, nr=122, rings=13,
                     rfov=93.0, zfov=34.459, rscanner=400.0, span=5, mrd=8),
         "c2": dict(nx=256, ny=256, nz=81, dx=2.0445, dz=2.027, nproj=252, nr=344, rings=64,
                    rfov=261.696, zfov=164.187, rscanner=400.0, span=11, mrd=60)}[geom]
    log(f"case: fused backprojection update, {geom}, strict={strict}")
    p = plan3d(ctx, g)
    rng = np.random.default_rng(7)
    S, s = 21, 3
    img = phantoms.cylinder3d(g["nx"], g["nz"], g["dx"], g["dz"], radius=0.6 * g["nx"] * g["dx"] / 2)
    sino = p.project(dev(torch, img), subset=s, nsubsets=S)
    sens = p.backproject(torch.ones_like(sino), subset=s, nsubsets=S, mode=hip.BP_GATHER)
    thr = hip.mlem_threshold(ctx, sens)
    x0 = dev(torch, rng.uniform(0.5, 2.0, img.shape))
    b = torch.zeros_like(x0)
    p.backproject(sino, out=b, subset=s, nsubsets=S, mode=hip.BP_GATHER)
    ref = hip.em_update(ctx, x0.clone(), b, sens, thr, strict)
    got = p.backproject_update(sino, x0.clone(), sens, thr, strict, subset=s, nsubsets=S)
    torch.cuda.synchronize()
    r, o = host(ref), host(got)
    assert (o == 0).sum() > 0 and (o > 0).sum() > 0  # both branches of the rule taken
    assert np.array_equal(o.view(np.int32), r.view(np.int32))
