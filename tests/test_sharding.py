"""Host side of the sharded 3D job: the ranks' sinogram sets partition the sinogram,
each segment spread evenly over the ranks."""
import numpy as np
import pytest

from apirl_amd import sharding


@pytest.mark.parametrize("layout", sharding.LAYOUTS)
@pytest.mark.parametrize("nsino,nranks", [(837, 8), (73, 4), (5, 8), (1, 1)])
def test_shards_partition_the_sinogram(nsino, nranks, layout):
    parts = [sharding.shard_sinograms(nsino, nranks, q, layout) for q in range(nranks)]
    allidx = np.concatenate(parts)
    assert sorted(allidx.tolist()) == list(range(nsino))
    sizes = [len(p) for p in parts]
    assert max(sizes) - min(sizes) <= 1
    step = nranks if layout == "interleaved" else 1
    for p in parts:
        assert np.all(np.diff(p) == step)


def test_balanced_shards_equalise_weight():
    """Contiguous runs whose weights differ by at most one sinogram's weight, on the
    256x256x81 job's ring table (837 sinograms, obliquity grows along the list)."""
    from apirl_amd import hip
    z1, z2 = hip.ring_table(64, 164.187, 11, 60)
    w = sharding.sinogram_weights(z1, z2, 256 * 2.0445)
    parts = [sharding.shard_sinograms(len(z1), 8, q, "balanced", w) for q in range(8)]
    assert sorted(np.concatenate(parts).tolist()) == list(range(len(z1)))
    tot = [w[p].sum() for p in parts]
    assert max(tot) - min(tot) <= 2 * w.max()
    assert len(parts[0]) > len(parts[-1])  # the direct planes cost less per sinogram
    for p in parts:
        assert np.all(np.diff(p) == 1)


def test_rank_out_of_range_raises():
    with pytest.raises(ValueError):
        sharding.shard_sinograms(10, 4, 4)
    with pytest.raises(ValueError):
        sharding.shard_sinograms(10, 4, -1)
    with pytest.raises(ValueError):
        sharding.shard_sinograms(10, 4, 0, "rows")
