"""A 3D job sharded by sinograms (apirl_amd.sharding): each rank's plan holds a
contiguous run of the full sinogram's sinograms (of equal length or of about equal
cost), or every n-th one.  Its projection equals those rows of
the full projection bit for bit, and the rank-ordered sum of the ranks' partial
backprojections and sensitivities (the all-reduce of the multi-GPU design) equals
the whole job's within the per-voxel bar.  One OSEM iteration run rank by rank,
with the partial images summed before each update, meets the image bar against the
whole job's."""
import json

import numpy as np
import pytest

from apirl_amd import hip, sharding
from test_gpu_parity import (IMG_RL2, assert_bitexact, assert_voxel_close, ctx, dev,  # noqa: F401
                             host, log, plan3d, rel_l2, torch)

pytestmark = pytest.mark.gpu

NRANKS = 4
NSUB = 7


def geometry(golden):
    f = golden("mid3d_project_backproject.npz")
    return f, json.loads(str(f["geom"]))


def shards(ctx, g, layout):
    return [sharding.shard_plan(ctx, g["nx"], g["ny"], g["nz"], g["dx"], g["dz"], g["nproj"],
                                g["nr"], g["rfov"], g["rscanner"], g["rings"], g["zfov"],
                                g["span"], g["mrd"], NRANKS, q, layout) for q in range(NRANKS)]


@pytest.mark.parametrize("layout", sharding.LAYOUTS)
def test_shard_projection_and_backprojection_sum(torch, ctx, golden, layout):
    log(f"case: mid3d sharded over {NRANKS} ranks by sinograms, {layout}")
    f, g = geometry(golden)
    full = plan3d(ctx, g)
    ranks = shards(ctx, g, layout)
    assert sorted(np.concatenate([i for _, i in ranks]).tolist()) == list(range(full.nsino))
    img = dev(torch, f["img"])
    whole = host(full.project(img))
    rng = np.random.default_rng(41)
    sino = rng.random(full.subset_shape(2, NSUB), dtype=np.float32)
    sino[rng.random(sino.shape) < 0.1] = 0
    bp_full = host(full.backproject(dev(torch, sino), subset=2, nsubsets=NSUB,
                                    mode=hip.BP_GATHER))
    sens_full = host(full.sensitivity(subset=2, nsubsets=NSUB, mode=hip.BP_GATHER,
                                      force_positive=0))
    bp_sum = np.zeros_like(bp_full)
    sens_sum = np.zeros_like(sens_full)
    for plan, idx in ranks:
        assert_bitexact(host(plan.project(img)), whole[idx])
        bp_sum += host(plan.backproject(dev(torch, np.ascontiguousarray(sino[idx])), subset=2,
                                        nsubsets=NSUB, mode=hip.BP_GATHER))
        sens_sum += host(plan.sensitivity(subset=2, nsubsets=NSUB, mode=hip.BP_GATHER,
                                          force_positive=0))
    assert_voxel_close(bp_sum, bp_full)
    assert_voxel_close(sens_sum, sens_full)


def test_sharded_osem_iteration_matches_whole_job(torch, ctx, golden):
    log(f"case: one OSEM iteration ({NSUB} subsets), ranks summed before each update")
    f, g = geometry(golden)
    full = plan3d(ctx, g)
    ranks = shards(ctx, g, "balanced")
    meas = f["meas"].astype(np.float32).reshape(full.subset_shape())
    init = torch.ones(full.image_shape(), dtype=torch.float32, device="cuda")
    ref, _ = full.reconstruct(dev(torch, meas), init, 1, subsets=NSUB, bp_mode=hip.BP_GATHER)
    x = torch.ones(full.image_shape(), dtype=torch.float32, device="cuda")
    for s in range(NSUB):
        sens = torch.zeros(full.image_shape(), dtype=torch.float32, device="cuda")
        bp = torch.zeros(full.image_shape(), dtype=torch.float32, device="cuda")
        for plan, idx in ranks:  # rank order: the all-reduce's fixed order
            m = dev(torch, np.ascontiguousarray(meas[idx]))
            ratio = plan.project_ratio(x, m, subset=s, nsubsets=NSUB, want_lik=False)[0]
            part = plan.backproject(ratio, subset=s, nsubsets=NSUB, mode=hip.BP_GATHER)
            bp += part
            sens += plan.sensitivity(subset=s, nsubsets=NSUB, mode=hip.BP_GATHER,
                                     force_positive=0)
        # (forcePositive, which the whole job applies to its sensitivity, changes
        # nothing in a sum of non-negative terms)
        hip.em_update(full.ctx, x, bp, sens, 0.0, 0)
    torch.cuda.synchronize()
    got, want = host(x), host(ref)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert rel_l2(got, want) <= IMG_RL2
