"""The C-ABI boundary on the CPU: the library loads, exports every symbol the header
declares, and its host-side geometry helpers equal the oracle's (which is pinned to
the reference)."""
import os
import re

import numpy as np

import _oracle as O
from apirl_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "apirl_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(apirl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = hip.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(hip.SIGNATURES), set(syms) ^ set(hip.SIGNATURES)


def test_sino_grid_matches_oracle():
    for nang, nr, rfov in [(192, 192, 128.0), (252, 344, 261.696), (7, 13, 3.3)]:
        a1, r1 = hip.sino_grid(nang, nr, rfov)
        a2, r2 = O.sino_grid(nang, nr, rfov)
        assert np.array_equal(a1.view(np.int32), a2.view(np.int32))
        assert np.array_equal(r1.view(np.int32), r2.view(np.int32))


def test_ring_table_matches_oracle():
    for rings, zfov, span, mrd in [(64, 164.187, 11, 60), (13, 34.459, 5, 8), (24, 95.3, 5, 14)]:
        z1, z2 = hip.ring_table(rings, zfov, span, mrd)
        nsino, mn, mx = hip.segment_table(rings, span, mrd)
        za, zb = O.ring_table(rings, zfov, span, mrd)
        assert len(z1) == len(za) == int(nsino.sum())
        # the oracle returns the final P.Z = zmean - w/2 + w/2 + 0; recompute that here
        w = np.float32(zfov) / np.float32(rings)
        fa = ((z1 - w / np.float32(2)) + w / np.float32(2)) + np.float32(0) * w
        assert np.array_equal(fa.view(np.int32), za.view(np.int32))


def test_c2_sinogram_count():
    nsino, mn, mx = hip.segment_table(64, 11, 60)
    assert list(nsino) == [127, 115, 115, 93, 93, 71, 71, 49, 49, 27, 27]
    assert nsino.sum() == 837


def test_subset_rule():
    lib = hip.load()
    for nang, S in [(252, 21), (192, 7), (10, 3)]:
        counts = [lib.apirl_subset_num_angles(nang, S, s) for s in range(S)]
        assert sum(counts) == nang
        for s in range(S):
            assert counts[s] == len(O.subset_angles(nang, S, s))


def test_no_gpu_raises_instead_of_falling_back():
    import torch
    if torch.cuda.is_available():
        return
    try:
        hip.Context(0)
    except hip.ApirlError:
        return
    raise AssertionError("Context() must raise without a GPU")
