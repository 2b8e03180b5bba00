"""The allocation-free entry point apirl_project_ratio_noalloc, called through the
C-ABI, against apirl_project_ratio: the ratio sinogram bit for bit and the likelihood
equal, on two calls in a row (the second with a smaller input), without a change in
the device's free memory; a likelihood over more bins than its partials array holds
is refused before any launch."""
import ctypes

import numpy as np
import pytest

from apirl_amd import hip, phantoms

pytestmark = pytest.mark.gpu

MID = dict(nx=64, ny=64, nz=17, dx=2.0445, dz=2.027, nproj=42, nr=86, rings=13, rfov=65.424,
           zfov=34.459, rscanner=400.0, span=5, mrd=8)


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a ROCm device"
    return t


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def call(ctx, name, plan, img, meas, add, ratio, lik, subset, nsub, flags):
    st = getattr(ctx.lib, name)(plan.h, ptr(img), ptr(meas), ptr(add), ptr(ratio), ptr(lik),
                                subset, nsub, flags, hip.PREC_REFERENCE)
    return st


def test_noalloc_equals_project_ratio_two_calls(torch):
    ctx = hip.Context(0)  # a fresh context: its own partials buffer was never allocated
    g = MID
    p = hip.Plan.for_3d(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"], g["nr"],
                        g["rfov"], g["rscanner"], g["rings"], g["zfov"], g["span"], g["mrd"])
    img = torch.from_numpy(phantoms.cylinder3d(g["nx"], g["nz"], g["dx"], g["dz"],
                                               radius=40.0)).cuda()
    full = p.subset_shape()
    rng = np.random.default_rng(5)
    meas = torch.from_numpy(rng.poisson(3.0, full).astype(np.float32)).cuda()
    add, _ = phantoms.correction_sinograms(full[0], full[1], full[2])
    add = torch.from_numpy(add).cuda()
    # call 1: the whole sinogram with the additive term; call 2: subset 3 of 7, smaller
    calls = ((0, 1, 0), (3, 7, hip.RATIO_ADDITIVE_LOCAL))
    bufs = [(torch.full(p.subset_shape(s, n), -1.0, dtype=torch.float32, device="cuda"),
             torch.full((1,), -1.0, dtype=torch.float64, device="cuda")) for s, n, _ in calls]
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for k, ((subset, nsub, flags), (r_na, lik_na)) in enumerate(zip(calls, bufs)):
        log(f"case: noalloc call {k + 1}, subset {subset} of {nsub}")
        assert call(ctx, "apirl_project_ratio_noalloc", p, img, meas, add, r_na, lik_na,
                    subset, nsub, flags) == 0, ctx.lib.apirl_last_error(ctx.h)
    torch.cuda.synchronize()
    # nothing was allocated by the allocation-free calls
    assert torch.cuda.mem_get_info()[0] == free0
    outs = [(s, n, f, r.cpu().numpy(), float(lk.item())) for (s, n, f), (r, lk) in zip(calls, bufs)]
    for subset, nsub, flags, r_na, lik_na in outs:
        shape = p.subset_shape(subset, nsub)
        r = torch.empty(shape, dtype=torch.float32, device="cuda")
        lik = torch.zeros(1, dtype=torch.float64, device="cuda")
        assert call(ctx, "apirl_project_ratio", p, img, meas, add, r, lik, subset, nsub,
                    flags) == 0
        torch.cuda.synchronize()
        ref = r.cpu().numpy()
        assert np.array_equal(r_na.view(np.int32), ref.view(np.int32))
        assert lik_na == float(lik.item()) and np.isfinite(lik_na)
    # without the likelihood the partials are not touched
    r2 = torch.empty(p.subset_shape(3, 7), dtype=torch.float32, device="cuda")
    assert call(ctx, "apirl_project_ratio_noalloc", p, img, meas, None, r2, None, 3, 7, 0) == 0
    torch.cuda.synchronize()


def test_noalloc_refuses_likelihood_beyond_its_partials(torch):
    """A sinogram of more than APIRL_NOALLOC_MAX_BINS = 2^27 bins (1600 sinograms x 252
    angles x 344 bins) with a likelihood: APIRL_ERR_INVALID, nothing launched."""
    log("case: noalloc size limit")
    ctx = hip.Context(0)
    nsino = 1600
    z = np.full(nsino, 1.0, np.float32)
    p = hip.Plan(ctx, 8, 8, 2, 4.0, 1.0, 252, 344, 16.0, 400.0, dims=3, z1=z, z2=z,
                 ring_width=1.0)
    assert nsino * 252 * 344 > (1 << 27)
    img = torch.zeros(p.image_shape(), dtype=torch.float32, device="cuda")
    meas = torch.zeros(p.subset_shape(), dtype=torch.float32, device="cuda")
    ratio = torch.zeros(p.subset_shape(), dtype=torch.float32, device="cuda")
    lik = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = call(ctx, "apirl_project_ratio_noalloc", p, img, meas, None, ratio, lik, 0, 1, 0)
    assert st == 1, st  # APIRL_ERR_INVALID
    assert b"APIRL_NOALLOC_MAX_BINS" in ctx.lib.apirl_last_error(ctx.h)
