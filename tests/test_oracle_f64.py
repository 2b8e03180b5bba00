"""Pins the FP64 oracle (the reference's 3D Siddon carried out in double,
oracle/apirl_oracle.c) against an independent double-precision ray tracer: the
textbook form of Siddon's method (all plane crossings inside the FOV merged and
sorted, each interval assigned to the voxel of its midpoint), written here in numpy.

The two share nothing but the voxel grid of the reference (recon/src/Siddon.cpp:
x, y faces at -R + i*dx with R = nx*dx/2; axial faces at dz/2 + k*dz, :162).  For
every LOR the oracle's list must carry the brute-force chord of each voxel it
names (to 1e-9 mm), and the only voxel with a chord that the list leaves out is the
exit voxel (the reference's loop stops after |di|+|dj|+|dk| steps, Siddon.cpp:220).
That holds on LORs lying on a voxel face too, where the float traversal may put the
whole chord on the other side."""
import numpy as np
import pytest

import _oracle as O

NX, NZ, DX, DZ = 8, 5, 2.5, 2.0


def brute_force(p, v, nx, nz, dx, dz):
    p = np.asarray(p, np.float64)
    v = np.asarray(v, np.float64)
    R = nx * dx / 2.0
    L = float(np.sqrt(np.sum(v * v)))
    a0, a1 = -np.inf, np.inf
    for d in range(2):  # the reference clips to the square FOV only (Siddon.cpp:83-145)
        if v[d] == 0:
            continue
        t0, t1 = sorted(((-R - p[d]) / v[d], (R - p[d]) / v[d]))
        a0, a1 = max(a0, t0), min(a1, t1)
    a0 = 0.0 if (a0 < 0 or a0 > 1) else a0
    a1 = 1.0 if (a1 > 1 or a1 < 0) else a1
    alphas = [a0, a1]
    for d, (origin, step, n) in enumerate([(-R, dx, nx), (-R, dx, nx), (dz / 2.0, dz, None)]):
        if v[d] == 0:
            continue
        if n is None:
            z0, z1 = sorted((p[2] + a0 * v[2], p[2] + a1 * v[2]))
            ks = range(int(np.floor((z0 - origin) / step)) - 1,
                       int(np.ceil((z1 - origin) / step)) + 2)
        else:
            ks = range(n + 1)
        for k in ks:
            a = (origin + k * step - p[d]) / v[d]
            if a0 < a < a1:
                alphas.append(a)
    alphas = np.sort(np.array(alphas))
    out = {}
    for lo, hi in zip(alphas[:-1], alphas[1:]):
        if hi <= lo:
            continue
        m = p + 0.5 * (lo + hi) * v
        key = (int(np.floor((m[0] + R) / dx)), int(np.floor((m[1] + R) / dx)),
               int(np.floor((m[2] - dz / 2.0) / dz)))
        out[key] = out.get(key, 0.0) + (hi - lo) * L
    return out, (a1 - a0) * L


def rays(seed=0, n=300):
    rng = np.random.default_rng(seed)
    R = NX * DX / 2.0
    H = NZ * DZ
    out = []
    for _ in range(n):
        c = np.array([rng.uniform(-R, R), rng.uniform(-R, R), rng.uniform(DZ, H)])
        phi = rng.uniform(0, np.pi)
        tz = rng.uniform(-0.3, 0.3)
        d = np.array([np.cos(phi), np.sin(phi), tz])
        out.append((c - 400 * d, 800 * d))
    # axis-parallel, diagonal and flat rays
    out.append((np.array([-400.0, 1.3, 5.1]), np.array([800.0, 0.0, 0.0])))
    out.append((np.array([2.2, -400.0, 4.3]), np.array([0.0, 800.0, 1.0])))
    out.append((np.array([-400.0, -400.0, 6.7]), np.array([800.0, 800.0, 0.0])))
    out.append((np.array([-400.0, 3.9, 2.5]), np.array([800.0, -7.0, 5.0])))
    return [(np.float32(p), np.float32(v)) for p, v in out]


@pytest.fixture(scope="module")
def im():
    return O.image_desc(NX, NX, NZ, DX, DZ)


def test_f64_oracle_chords_match_independent_tracer(im):
    checked = 0
    for p, v in rays():
        lst = O.siddon3d_f64(p, v, im)
        bf, total = brute_force(p.astype(np.float64), v.astype(np.float64), NX, NZ, DX, DZ)
        keys = set()
        for i, j, k, w in lst:
            assert abs(w - bf.get((i, j, k), 0.0)) <= 1e-9 * DX, (p, v, (i, j, k), w)
            keys.add((i, j, k))
        missing = {key: w for key, w in bf.items() if w > 1e-9 and key not in keys}
        assert len(missing) <= 1, (p, v, missing)
        if lst:
            assert abs(sum(w for *_, w in lst) + sum(missing.values()) - total) <= 1e-9 * total
            checked += 1
    assert checked > 100


def test_f64_oracle_is_close_to_the_float_weights(im):
    """Where the float and the double traversal differ.  The list length
    |di|+|dj|+|dk| comes from the exit voxel, and LORs leave the square FOV exactly on
    its face x or y = +-R, where floor() of the exit coordinate flips between float
    and double: the lists then differ in their last voxel.  Apart from the last voxel
    every voxel gets the same length up to float rounding of alpha (~1e-7 of the LOR
    length; near a tie of two crossings the two may step through different corner
    voxels, with segments of that size)."""
    worst, same = 0.0, 0
    rs = rays(seed=1, n=100)
    for p, v in rs:
        fl = O.siddon3d(p, v, im)
        dl = O.siddon3d_f64(p, v, im)
        same += [s[:3] for s in fl] == [s[:3] for s in dl]
        n = min(len(fl), len(dl)) - 1
        f, d = {}, {}
        for i, j, k, w in fl[:n]:
            f[(i, j, k)] = f.get((i, j, k), 0.0) + w
        for i, j, k, w in dl[:n]:
            d[(i, j, k)] = d.get((i, j, k), 0.0) + w
        for key in set(f) | set(d):
            worst = max(worst, abs(f.get(key, 0.0) - d.get(key, 0.0)))
    assert worst < 1e-3
    assert same >= 0.5 * len(rs)


def test_planted_wrong_axial_offset_is_caught(im):
    """A tracer with the axial grid at k*dz (no dz/2 offset) disagrees with the oracle:
    the comparison above can see a mistake of that kind."""
    p, v = rays(seed=2, n=1)[0]
    lst = O.siddon3d_f64(p, v, im)
    bf, _ = brute_force(p.astype(np.float64), v.astype(np.float64) , NX, NZ, DX, DZ)
    bf_wrong, _ = brute_force(p.astype(np.float64) + np.array([0, 0, DZ / 2.0]),
                              v.astype(np.float64), NX, NZ, DX, DZ)
    good = all(abs(w - bf.get((i, j, k), 0.0)) <= 1e-9 for i, j, k, w in lst)
    bad = all(abs(w - bf_wrong.get((i, j, k), 0.0)) <= 1e-9 for i, j, k, w in lst)
    assert good and not bad


MINI = dict(nx=16, ny=16, nz=7, dx=4.0, dz=4.0, nproj=12, nr=24, rings=4, rfov=46.0,
            zfov=28.0, span=3, mrd=2, rscanner=120.0)


def mini_geometry():
    g = MINI
    im = O.image_desc(g["nx"], g["ny"], g["nz"], g["dx"], g["dz"])
    ang, r = O.sino_grid(g["nproj"], g["nr"], g["rfov"])
    za, zb = O.ring_table(g["rings"], g["zfov"], g["span"], g["mrd"])
    return g, im, ang, r, za, zb


def test_f64_sampled_bins_equal_full_f64_projection():
    g, im, ang, r, za, zb = mini_geometry()
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 2, (g["nz"], g["ny"], g["nx"]))
    full = O.project3d_f64(im, img, ang, r, g["rscanner"], za, zb)
    full = full.reshape(len(za), g["nproj"], g["nr"])
    s = rng.integers(0, len(za), 200)
    a = rng.integers(0, g["nproj"], 200)
    q = rng.integers(0, g["nr"], 200)
    got = O.project3d_bins_f64(im, img, ang, r, g["rscanner"], za, zb, s, a, q)
    assert np.array_equal(got, full[s, a, q])


def test_f64_osem_loop_matches_numpy_over_the_f64_projectors():
    """The double OSEM driver against OSEM written out in numpy over the FP64
    projector and backprojector (pinned above): per-subset sensitivities, ratio rule
    (Sinogram3D.cpp:830-856), update with threshold 0 and >= (OsemSinogram3d.cpp:
    217-237), subsets in index order."""
    g, im, ang, r, za, zb = mini_geometry()
    rng = np.random.default_rng(4)
    x = rng.uniform(0.5, 2, (g["nz"], g["ny"], g["nx"])).astype(np.float32)
    meas = O.project3d(im, x, ang, r, g["rscanner"], za, zb).astype(np.float32)
    init = np.ones_like(x)
    S, T = 3, 2
    got = O.osem3d_f64(im, g["nproj"], g["nr"], g["rfov"], g["rscanner"], za, zb, meas, init,
                       S, T)
    img = init.astype(np.float64)
    subsets = [O.subset_angles(g["nproj"], S, s) for s in range(S)]
    sens = [O.backproject3d_f64(im, np.ones((len(za), len(a), g["nr"])), ang[a], r,
                                g["rscanner"], za, zb) for a in subsets]
    with np.errstate(divide="ignore", invalid="ignore"):
        for _ in range(T):
            for s, a in enumerate(subsets):
                y = meas[:, a, :].astype(np.float64)
                est = O.project3d_f64(im, img, ang[a], r, g["rscanner"], za, zb)
                ratio = np.where((y != 0) & (est != 0), y / est, 0.0)
                bp = O.backproject3d_f64(im, ratio, ang[a], r, g["rscanner"], za, zb)
                img = np.where(sens[s] >= 0, img * bp / sens[s], 0.0)
    assert np.array_equal(np.isnan(got), np.isnan(img))
    ok = ~np.isnan(img)
    assert np.allclose(got[ok], img[ok], rtol=1e-12, atol=0)


def test_face_aligned_lor_f64_follows_the_exact_geometry(im):
    """A LOR in the plane of an axial voxel face (z = dz/2 + k dz, in float): the FP64
    oracle puts its chord where the exact geometry of the float LOR does."""
    z = np.float32(DZ / 2 + 2 * DZ)
    p = np.array([-400.0, 1.3, z], np.float32)
    v = np.array([800.0, 3.0, 0.0], np.float32)
    lst = O.siddon3d_f64(p, v, im)
    bf, total = brute_force(p.astype(np.float64), v.astype(np.float64), NX, NZ, DX, DZ)
    for i, j, k, w in lst:
        assert abs(w - bf.get((i, j, k), 0.0)) <= 1e-9 * DX
    assert abs(sum(w for *_, w in lst) + max(bf.values()) - total) < total
