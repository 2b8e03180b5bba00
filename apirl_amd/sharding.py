"""One rank's share of a 3D job sharded by sinograms.

A single large 3D reconstruction shards over n GPUs by rays: rank q takes a share of
the sinograms of every subset (a contiguous run of them of about equal cost, or every
n-th one), with the image replicated.  Forward projection
needs no exchange; each rank's backprojection is a partial image, and the partial
images (and, once, the partial sensitivities) are summed across ranks before the
update.  This module builds a rank's plan on one GPU: the same image grid and
sinogram grid with the rank's sinograms only, so the kernels run on one GPU's real
share of the job."""
import numpy as np

from . import hip


LAYOUTS = ("balanced", "block", "interleaved")

# Cost of a sinogram relative to one of the direct plane, per unit of axial ring
# distance |z2 - z1| over the image width: the gather backprojector's time per
# sinogram grows with obliquity (an oblique LOR reaches more sinograms of a slab).
# Calibrated on the 256x256x81 job (523 mm wide) from the per-rank times of the block
# and balanced layouts: about 1.9x at |z2 - z1| = 141 mm.
OBLIQUE_COST = 3.5


def shard_sinograms(nsino, nranks, rank, layout="balanced", weights=None):
    """Indices of rank's sinograms in the full sinogram's order.
    block: the contiguous run [rank nsino / nranks, (rank + 1) nsino / nranks), so a
    rank holds one or two segments and its backprojection scans only those;
    balanced: contiguous runs of equal total weight (weights: one per sinogram);
    interleaved: the sinograms rank, rank + nranks, ..., every segment spread over
    all ranks."""
    if not (0 <= rank < nranks):
        raise ValueError("rank must be in [0, nranks)")
    if layout == "interleaved":
        return np.arange(rank, nsino, nranks)
    if layout == "block":
        return np.arange(rank * nsino // nranks, (rank + 1) * nsino // nranks)
    if layout == "balanced":
        w = np.ones(nsino) if weights is None else np.asarray(weights, np.float64)
        if w.shape != (nsino,) or np.any(w <= 0):
            raise ValueError("balanced layout needs one positive weight per sinogram")
        # sinogram s goes to the rank whose share of the total weight holds the
        # middle of s's weight
        mid = np.cumsum(w) - 0.5 * w
        owner = np.minimum((mid * nranks / w.sum()).astype(np.int64), nranks - 1)
        return np.flatnonzero(owner == rank)
    raise ValueError(f"layout must be one of {LAYOUTS}")


def sinogram_weights(z1, z2, image_width_mm):
    """Relative backprojection cost of each sinogram (see OBLIQUE_COST)."""
    dz = np.abs(np.asarray(z2, np.float64) - np.asarray(z1, np.float64))
    return 1.0 + OBLIQUE_COST * dz / image_width_mm


def shard_plan(ctx, nx, ny, nz, dx, dz, nang, nr, rfov, rscanner, rings, zfov, span, mrd,
               nranks, rank, layout="balanced"):
    """The plan of rank's share of a cylindrical 3D sinogram (segment table as
    hip.Plan.for_3d builds it); returns (plan, sinogram indices)."""
    z1, z2 = hip.ring_table(rings, zfov, span, mrd)
    idx = shard_sinograms(len(z1), nranks, rank, layout,
                          sinogram_weights(z1, z2, nx * dx) if layout == "balanced" else None)
    plan = hip.Plan(ctx, nx, ny, nz, dx, dz, nang, nr, rfov, rscanner, dims=3, z1=z1[idx],
                    z2=z2[idx], ring_width=np.float32(zfov) / np.float32(rings))
    return plan, idx
