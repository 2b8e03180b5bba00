"""Slice-by-slice 2D MLEM (recon/src/Mlem2dMultiple.cpp:64-84), run as one batch.

The reference reconstructs each 2D sinogram of a multi-slice file
(Sinograms2DinCylindrical3Dpet) with its own Mlem2d, one after another: every slice
starts from an image of ones (the initial estimate only gives the size), the radial
FOV is the image's, (dx * nx) / 2 in float (Mlem2d.cpp:12), the multiplicative and
additive sinograms are taken slice by slice, and the slices are stacked into the
volume.  Here every slice is a job of one apirl_batch on a shared 2D plan, so the
slices run together; each slice's image equals the single-job driver's, bit for bit
(gather backprojector, reference arithmetic)."""
import numpy as np

from . import hip


def fov_radius(nx, dx):
    """Image FOV radius as the reference computes it: (dx * nx) / 2 in float
    (data/src/Images.cpp:378)."""
    return float(np.float32(np.float32(dx) * np.float32(nx)) / np.float32(2))


def slice_plan(ctx, nx, dx, nproj, nr, rscanner, rfov=None):
    """The 2D plan every slice uses: the radial FOV is the image's (the Mlem2d
    override) unless given (the sinograms' own, as project / backproject read it)."""
    return hip.Plan.for_2d(ctx, nx, dx, nproj, nr, fov_radius(nx, dx) if rfov is None else rfov,
                           rscanner)


def mlem2d_multiple(plan, sinos, iterations, init=None, mult=None, additive=None,
                    max_jobs=64):
    """sinos: device tensor [nslices, nproj, nr]; init: None (ones, as the reference)
    or [nslices, nx, nx] to continue from; mult / additive: None or the shape of sinos.
    Returns the volume [nslices, nx, nx]."""
    import torch

    if sinos.dim() != 3 or tuple(sinos.shape[1:]) != (plan.nang, plan.nr):
        raise ValueError("sinos must be [slices, angles, radial bins] of the plan")
    nz = sinos.shape[0]
    for name, t in (("mult", mult), ("additive", additive)):
        if t is not None and tuple(t.shape) != tuple(sinos.shape):
            raise ValueError(f"{name} must have the shape of the sinograms")
    nx = plan.nx
    if init is None:
        ones = torch.ones((nx, nx), dtype=torch.float32, device=sinos.device)
        first = [ones] * nz
    else:
        if tuple(init.shape) != (nz, nx, nx):
            raise ValueError("init must be [slices, nx, nx]")
        first = [init[k] for k in range(nz)]
    vol = torch.empty((nz, nx, nx), dtype=torch.float32, device=sinos.device)
    cap = max(1, min(nz, max_jobs))
    batch = hip.Batch(plan.ctx, max_jobs=cap, max_sinograms=1)
    try:
        pending = list(range(nz))
        running = {}
        while True:
            while pending and len(running) < cap:
                k = pending.pop(0)
                running[k] = batch.submit(
                    plan, sinos[k], first[k], vol[k], iterations, 1, osem_rules=0,
                    additive=None if additive is None else additive[k],
                    mult=None if mult is None else mult[k])
            active = batch.step()
            for k in list(running):
                if batch.done(running[k]):
                    batch.release(running.pop(k))
            if active == 0 and not pending:
                break
        batch.synchronize()
    finally:
        batch.close()
    return vol


def sensitivities(plan, nz, mult=None, out=None):
    """Each slice's sensitivity image as its Mlem2d computes it (A^T 1 or A^T m, no
    forcePositive, Mlem2d.cpp:262-278)."""
    import torch

    if out is None:
        out = torch.empty((nz, plan.nx, plan.nx), dtype=torch.float32, device="cuda")
    for k in range(nz):
        plan.sensitivity(out=out[k], mult=None if mult is None else mult[k],
                         mode=hip.BP_GATHER, force_positive=0)
    return out
