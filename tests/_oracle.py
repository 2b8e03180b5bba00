"""ctypes binding of the oracle (oracle/liboracle.so): the CPU restatement of the
reference's Siddon path.  Test infrastructure only -- imported by tests/, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``, always as
the checker, never as the thing measured or shipped."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

_f = ctypes.c_float
_i = ctypes.c_int
_l = ctypes.c_long
_P = ctypes.c_void_p


class OrcImage(ctypes.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("dx", _f), ("dy", _f), ("dz", _f)]


class OrcSeg(ctypes.Structure):
    _fields_ = [("i", _i), ("j", _i), ("k", _i), ("w", _f)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            # built by __graft_entry__.build() (make -C oracle); never built here, where
            # the process may already have initialised the GPU and must not fork a build
            raise RuntimeError(f"{LIB_PATH} is missing: run `make -C oracle` (or "
                               "__graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        P = _P
        sig = {
            "orc_sino_grid": (None, [_i, _i, _f, P, P]),
            "orc_lor_points": (_i, [_f, _f, _f, _f, P, P, P, P]),
            "orc_ring_table": (_i, [_i, _f, _i, P, P, P, P, P]),
            "orc_siddon2d": (_i, [_f, _f, _f, _f, P, P, _i]),
            "orc_siddon3d": (_i, [_f, _f, _f, _f, _f, _f, P, P, _i]),
            "orc_project2d": (None, [P, P, _i, P, _i, P, _f, _f, P]),
            "orc_project3d": (None, [P, P, _i, P, _i, P, _f, _i, P, P, P]),
            "orc_backproject2d": (None, [P, P, _i, P, _i, P, _f, _f, P]),
            "orc_divide_backproject2d": (None, [P, P, P, _i, P, _i, P, _f, _f, P]),
            "orc_subsample_r": (None, [_i, P, _i, P]),
            "orc_mmr_lor_table": (None, [_i, _i, P]),
            "orc_project3d_pts": (None, [P, P, _i, _i, P, _i, P, P, P]),
            "orc_backproject3d_pts": (None, [P, P, _i, _i, P, _i, P, P, P]),
            "orc_ring_table_os": (_i, [_i, _f, _i, P, P, P, _i, P, P]),
            "orc_project3d_os": (None, [P, P, _i, P, _i, P, _i, _f, _i, P, P, _i, P]),
            "orc_backproject3d_os": (None, [P, P, _i, P, _i, P, _i, _f, _i, P, P, _i, P]),
            "orc_backproject3d": (None, [P, P, _i, P, _i, P, _f, _i, P, P, P]),
            "orc_divide_backproject3d": (None, [P, P, P, _i, P, _i, P, _f, _i, P, P, P]),
            "orc_inverse_divide": (None, [P, P, _l]),
            "orc_likelihood": (_f, [P, P, _l, _i]),
            "orc_mlem_threshold": (_f, [P, _l]),
            "orc_min_nonzero": (_f, [P, _l]),
            "orc_em_update": (None, [P, P, P, _l, _f, _i]),
            "orc_subset_angles": (_i, [_i, _i, _i, P]),
            "orc_mlem2d": (_i, [P, _i, _i, _f, P, P, _i, P, P, P]),
            "orc_mlem3d": (_i, [P, _i, _i, _f, _f, _i, P, P, P, P, _i, P, P, P]),
            "orc_mlem3d_pts": (_i, [P, _i, _i, P, _i, P, P, P, P, _i, P, P, P]),
            "orc_osem3d": (_i, [P, _i, _i, _f, _f, _i, P, P, P, P, _i, _i, P, P, P]),
            "orc_osem3d_corr": (_i, [P, _i, _i, _f, _f, _i, P, P, P, P, _i, _i, P, P, _i, P,
                                     P, P]),
            "orc_osem3d_pts_corr": (_i, [P, _i, _i, P, P, _i, P, P, P, P, _i, _i, P, P, _i,
                                         P, P, P]),
            "orc_mlem3d_pts_corr": (_i, [P, _i, _i, P, _i, P, P, P, P, _i, P, P, P, P, P]),
            "orc_mmr_subset_lor_table": (None, [_i, _i, P]),
            "orc_siddon3d_f64": (_i, [P, P, P, P, P, _i]),
            "orc_project3d_bins_f64": (None, [P, P, P, P, _f, P, P, _l, P, P, P, P]),
            "orc_osem3d_f64": (_i, [P, _i, _i, _f, _f, _i, P, P, P, P, _i, _i, P]),
            "orc_project3d_f64": (None, [P, P, _i, P, _i, P, _f, _i, P, P, P]),
            "orc_backproject3d_f64": (None, [P, P, _i, P, _i, P, _f, _i, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def image_desc(nx, ny, nz, dx, dz=1.0):
    return OrcImage(nx, ny, nz, dx, dx, dz)


def sino_grid(nang, nr, rfov):
    ang = np.zeros(nang, np.float32)
    r = np.zeros(nr, np.float32)
    lib().orc_sino_grid(nang, nr, rfov, _p(ang), _p(r))
    return ang, r


def seg_table(rings, span, mrd):
    """Segment table ordered 0, +1, -1, +2, -2 ... (min/max ring difference per segment)."""
    h = (span - 1) // 2
    nsino, mn, mx = [2 * rings - 1], [-h], [h]
    lo = h + 1
    while lo <= mrd:
        hi = min(lo + span - 1, mrd)
        n = sum(1 for s in range(2 * rings - 1)
                if any(0 <= s - b < rings and lo <= s - 2 * b <= hi for b in range(rings)))
        nsino += [n, n]
        mn += [lo, -hi]
        mx += [hi, -lo]
        lo += span
    return (np.array(nsino, np.int32), np.array(mn, np.int32), np.array(mx, np.int32))


def ring_table(rings, zfov, span, mrd):
    nsino, mn, mx = seg_table(rings, span, mrd)
    tot = int(nsino.sum())
    za = np.zeros(tot, np.float32)
    zb = np.zeros(tot, np.float32)
    n = lib().orc_ring_table(rings, zfov, len(nsino), _p(nsino), _p(mn), _p(mx), _p(za), _p(zb))
    assert n == tot, (n, tot)
    return za, zb


def mmr_lor_table(nang, nr):
    """Siemens mMR LOR endpoints (x1, y1, x2, y2) per bin, [nang, nr, 4]."""
    out = np.zeros((nang, nr, 4), np.float32)
    lib().orc_mmr_lor_table(nang, nr, _p(out))
    return out


def mmr_subset_lor_table(nang, nr):
    """LOR endpoints of the reference's mMR subset sinograms (cylindrical on 328 mm at
    the mMR's grids), [nang, nr, 4] over the full angle grid."""
    out = np.zeros((nang, nr, 4), np.float32)
    lib().orc_mmr_subset_lor_table(nang, nr, _p(out))
    return out


def project3d_pts(im, img, pts, za, zb):
    img = f32(img)
    nang, nr = pts.shape[:2]
    out = np.zeros((za.size, nang, nr), np.float32)
    lib().orc_project3d_pts(ctypes.byref(im), _p(img), nang, nr, _p(f32(pts)), za.size,
                            _p(f32(za)), _p(f32(zb)), _p(out))
    return out


def backproject3d_pts(im, sino, pts, za, zb):
    sino = f32(sino)
    nang, nr = pts.shape[:2]
    img = np.zeros((im.nz, im.ny, im.nx), np.float32)
    lib().orc_backproject3d_pts(ctypes.byref(im), _p(sino), nang, nr, _p(f32(pts)), za.size,
                                _p(f32(za)), _p(f32(zb)), _p(img))
    return img


def subsample_r(r, ns):
    """Radial positions of the ns sub-LORs of every bin (oversampled 3D Siddon)."""
    r = f32(r)
    out = np.zeros(r.size * ns, np.float32)
    lib().orc_subsample_r(r.size, _p(r), ns, _p(out))
    return out


def ring_table_os(rings, zfov, span, mrd, nas):
    nsino, mn, mx = seg_table(rings, span, mrd)
    tot = int(nsino.sum())
    za = np.zeros(tot * nas, np.float32)
    zb = np.zeros(tot * nas, np.float32)
    n = lib().orc_ring_table_os(rings, zfov, len(nsino), _p(nsino), _p(mn), _p(mx), nas, _p(za),
                                _p(zb))
    assert n == tot, (n, tot)
    return za, zb


def project3d_os(im, img, ang, r_sub, ns, rscanner, za, zb, nas):
    img = f32(img)
    nr = r_sub.size // ns
    nsino = za.size // nas
    out = np.zeros((nsino, len(ang), nr), np.float32)
    lib().orc_project3d_os(ctypes.byref(im), _p(img), len(ang), _p(f32(ang)), nr, _p(f32(r_sub)),
                           ns, rscanner, nsino, _p(f32(za)), _p(f32(zb)), nas, _p(out))
    return out


def backproject3d_os(im, sino, ang, r_sub, ns, rscanner, za, zb, nas):
    sino = f32(sino)
    nr = r_sub.size // ns
    nsino = za.size // nas
    img = np.zeros((im.nz, im.ny, im.nx), np.float32)
    lib().orc_backproject3d_os(ctypes.byref(im), _p(sino), len(ang), _p(f32(ang)), nr,
                               _p(f32(r_sub)), ns, rscanner, nsino, _p(f32(za)), _p(f32(zb)), nas,
                               _p(img))
    return img


def siddon2d(p0, v, im, cap=4096):
    out = (OrcSeg * cap)()
    n = lib().orc_siddon2d(p0[0], p0[1], v[0], v[1], ctypes.byref(im), out, cap)
    return [(out[m].i, out[m].j, out[m].w) for m in range(min(n, cap))]


def siddon3d(p0, v, im, cap=8192):
    out = (OrcSeg * cap)()
    n = lib().orc_siddon3d(p0[0], p0[1], p0[2], v[0], v[1], v[2], ctypes.byref(im), out, cap)
    return [(out[m].i, out[m].j, out[m].k, out[m].w) for m in range(min(n, cap))]


def siddon3d_f64(p0, v, im, cap=8192):
    """FP64 oracle list of one LOR: [(i, j, k, exact chord length in double)]."""
    out = (OrcSeg * cap)()
    w = np.zeros(cap, np.float64)
    p = f32(p0)
    vv = f32(v)
    n = lib().orc_siddon3d_f64(_p(p), _p(vv), ctypes.byref(im), out, _p(w), cap)
    return [(out[m].i, out[m].j, out[m].k, float(w[m])) for m in range(min(n, cap))]


def lor_points(ang, r, rscanner, rfov=float("inf")):
    a, b, c, d = _f(), _f(), _f(), _f()
    ok = lib().orc_lor_points(ang, r, rscanner, rfov, ctypes.byref(a), ctypes.byref(b),
                              ctypes.byref(c), ctypes.byref(d))
    return ok, (a.value, b.value), (c.value, d.value)


def project2d(im, img, ang, r, rfov, rscanner, sino=None):
    img = f32(img)
    if sino is None:
        sino = np.zeros((len(ang), len(r)), np.float32)
    lib().orc_project2d(ctypes.byref(im), _p(img), len(ang), _p(f32(ang)), len(r), _p(f32(r)),
                        rfov, rscanner, _p(sino))
    return sino


def backproject2d(im, sino, ang, r, rfov, rscanner, img=None):
    sino = f32(sino)
    if img is None:
        img = np.zeros((im.ny, im.nx), np.float32)
    lib().orc_backproject2d(ctypes.byref(im), _p(sino), len(ang), _p(f32(ang)), len(r),
                            _p(f32(r)), rfov, rscanner, _p(img))
    return img


def divide_backproject2d(im, meas, est, ang, r, rfov, rscanner, img=None):
    if img is None:
        img = np.zeros((im.ny, im.nx), np.float32)
    lib().orc_divide_backproject2d(ctypes.byref(im), _p(f32(meas)), _p(f32(est)), len(ang),
                                   _p(f32(ang)), len(r), _p(f32(r)), rfov, rscanner, _p(img))
    return img


def project3d(im, img, ang, r, rscanner, za, zb):
    img = f32(img)
    sino = np.zeros((len(za), len(ang), len(r)), np.float32)
    lib().orc_project3d(ctypes.byref(im), _p(img), len(ang), _p(f32(ang)), len(r), _p(f32(r)),
                        rscanner, len(za), _p(f32(za)), _p(f32(zb)), _p(sino))
    return sino


def backproject3d(im, sino, ang, r, rscanner, za, zb, img=None):
    sino = f32(sino)
    if img is None:
        img = np.zeros((im.nz, im.ny, im.nx), np.float32)
    lib().orc_backproject3d(ctypes.byref(im), _p(sino), len(ang), _p(f32(ang)), len(r),
                            _p(f32(r)), rscanner, len(za), _p(f32(za)), _p(f32(zb)), _p(img))
    return img


def divide_backproject3d(im, meas, est, ang, r, rscanner, za, zb, img=None):
    meas, est = f32(meas), f32(est)
    if img is None:
        img = np.zeros((im.nz, im.ny, im.nx), np.float32)
    lib().orc_divide_backproject3d(ctypes.byref(im), _p(meas), _p(est), len(ang), _p(f32(ang)),
                                   len(r), _p(f32(r)), rscanner, len(za), _p(f32(za)),
                                   _p(f32(zb)), _p(img))
    return img


def project3d_f64(im, img, ang, r, rscanner, za, zb):
    img = np.ascontiguousarray(img, np.float64)
    sino = np.zeros((len(za), len(ang), len(r)), np.float64)
    lib().orc_project3d_f64(ctypes.byref(im), _p(img), len(ang), _p(f32(ang)), len(r), _p(f32(r)),
                            rscanner, len(za), _p(f32(za)), _p(f32(zb)), _p(sino))
    return sino


def backproject3d_f64(im, sino, ang, r, rscanner, za, zb):
    sino = np.ascontiguousarray(sino, np.float64)
    img = np.zeros((im.nz, im.ny, im.nx), np.float64)
    lib().orc_backproject3d_f64(ctypes.byref(im), _p(sino), len(ang), _p(f32(ang)), len(r),
                                _p(f32(r)), rscanner, len(za), _p(f32(za)), _p(f32(zb)), _p(img))
    return img


def project3d_bins_f64(im, img, ang, r, rscanner, za, zb, s_idx, a_idx, r_idx):
    """FP64 oracle line integrals of the bins (s_idx, a_idx, r_idx) of the full sinogram."""
    img = np.ascontiguousarray(img, np.float64)
    si, ai, ri = (np.ascontiguousarray(x, np.int32) for x in (s_idx, a_idx, r_idx))
    out = np.zeros(len(si), np.float64)
    lib().orc_project3d_bins_f64(ctypes.byref(im), _p(img), _p(f32(ang)), _p(f32(r)), rscanner,
                                 _p(f32(za)), _p(f32(zb)), len(si), _p(si), _p(ai), _p(ri),
                                 _p(out))
    return out


def osem3d_f64(im, nang, nr, rfov, rscanner, za, zb, meas, init, nsubsets, iters):
    img = np.zeros((im.nz, im.ny, im.nx), np.float64)
    lib().orc_osem3d_f64(ctypes.byref(im), nang, nr, rfov, rscanner, len(za), _p(f32(za)),
                         _p(f32(zb)), _p(f32(meas)), _p(np.ascontiguousarray(init, np.float64)),
                         nsubsets, iters, _p(img))
    return img


def inverse_divide(est, meas):
    est = f32(est).copy()
    lib().orc_inverse_divide(_p(est), _p(f32(meas)), est.size)
    return est


def likelihood(est, meas, positive_only):
    return lib().orc_likelihood(_p(f32(est)), _p(f32(meas)), f32(est).size, int(positive_only))


def mlem_threshold(sens):
    return lib().orc_mlem_threshold(_p(f32(sens)), f32(sens).size)


def em_update(x, b, s, thr, strict):
    x = f32(x).copy()
    lib().orc_em_update(_p(x), _p(f32(b)), _p(f32(s)), x.size, thr, int(strict))
    return x


def subset_angles(nang, nsub, s):
    idx = np.zeros(nang, np.int32)
    n = lib().orc_subset_angles(nang, nsub, s, _p(idx))
    return idx[:n].copy()


def mlem2d(im, nang, nr, rscanner, meas, init, iters):
    npix = im.nx * im.ny
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix, np.float32)
    lik = np.zeros(iters + 1, np.float32)
    lib().orc_mlem2d(ctypes.byref(im), nang, nr, rscanner, _p(f32(meas)), _p(f32(init)), iters,
                     _p(img), _p(sens), _p(lik))
    return img.reshape(im.ny, im.nx), sens.reshape(im.ny, im.nx), lik


def mlem3d(im, nang, nr, rfov, rscanner, za, zb, meas, init, iters):
    npix = im.nx * im.ny * im.nz
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix, np.float32)
    lik = np.zeros(iters + 1, np.float32)
    lib().orc_mlem3d(ctypes.byref(im), nang, nr, rfov, rscanner, len(za), _p(f32(za)), _p(f32(zb)),
                     _p(f32(meas)), _p(f32(init)), iters, _p(img), _p(sens), _p(lik))
    shp = (im.nz, im.ny, im.nx)
    return img.reshape(shp), sens.reshape(shp), lik


def mlem3d_pts(im, pts, za, zb, meas, init, iters):
    """MlemSinogram3d over the LOR endpoints pts [nang, nr, 4] (e.g. mmr_lor_table)."""
    nang, nr = pts.shape[:2]
    npix = im.nx * im.ny * im.nz
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix, np.float32)
    lik = np.zeros(iters + 1, np.float32)
    lib().orc_mlem3d_pts(ctypes.byref(im), nang, nr, _p(f32(pts)), len(za), _p(f32(za)),
                         _p(f32(zb)), _p(f32(meas)), _p(f32(init)), iters, _p(img), _p(sens),
                         _p(lik))
    shp = (im.nz, im.ny, im.nx)
    return img.reshape(shp), sens.reshape(shp), lik


def osem3d(im, nang, nr, rfov, rscanner, za, zb, meas, init, nsub, iters):
    npix = im.nx * im.ny * im.nz
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix * nsub, np.float32)
    lik = np.zeros(iters * nsub + 1, np.float32)
    lib().orc_osem3d(ctypes.byref(im), nang, nr, rfov, rscanner, len(za), _p(f32(za)),
                     _p(f32(zb)), _p(f32(meas)), _p(f32(init)), nsub, iters, _p(img), _p(sens),
                     _p(lik))
    shp = (im.nz, im.ny, im.nx)
    return img.reshape(shp), sens.reshape((nsub,) + shp), lik


def osem3d_corr(im, nang, nr, rfov, rscanner, za, zb, meas, init, nsub, iters, add=None,
                mult=None, add_local=True):
    """OsemSinogram3d with an additive (no clamp; subset bin k takes the additive's angle
    k when add_local, as the reference does) and a multiplicative (sensitivity) term."""
    npix = im.nx * im.ny * im.nz
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix * nsub, np.float32)
    lik = np.zeros(iters * nsub + 1, np.float32)
    lib().orc_osem3d_corr(ctypes.byref(im), nang, nr, rfov, rscanner, len(za), _p(f32(za)),
                          _p(f32(zb)), _p(f32(meas)), _p(f32(init)), nsub, iters,
                          _p(f32(add)) if add is not None else None,
                          _p(f32(mult)) if mult is not None else None, 1 if add_local else 0,
                          _p(img), _p(sens), _p(lik))
    shp = (im.nz, im.ny, im.nx)
    return img.reshape(shp), sens.reshape((nsub,) + shp), lik


def osem3d_pts_corr(im, pts_sub, pts_full, za, zb, meas, init, nsub, iters, add=None,
                    mult=None, add_local=True):
    """osem3d_corr over LOR endpoints: subset bins on pts_sub, the final projection on
    pts_full (both [nang, nr, 4])."""
    nang, nr = pts_full.shape[:2]
    npix = im.nx * im.ny * im.nz
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix * nsub, np.float32)
    lik = np.zeros(iters * nsub + 1, np.float32)
    lib().orc_osem3d_pts_corr(ctypes.byref(im), nang, nr, _p(f32(pts_sub)), _p(f32(pts_full)),
                              len(za), _p(f32(za)), _p(f32(zb)), _p(f32(meas)), _p(f32(init)),
                              nsub, iters, _p(f32(add)) if add is not None else None,
                              _p(f32(mult)) if mult is not None else None,
                              1 if add_local else 0, _p(img), _p(sens), _p(lik))
    shp = (im.nz, im.ny, im.nx)
    return img.reshape(shp), sens.reshape((nsub,) + shp), lik


def mlem3d_pts_corr(im, pts, za, zb, meas, init, iters, add=None, mult=None):
    """MlemSinogram3d over LOR endpoints with the additive and multiplicative terms."""
    nang, nr = pts.shape[:2]
    npix = im.nx * im.ny * im.nz
    img = np.zeros(npix, np.float32)
    sens = np.zeros(npix, np.float32)
    lik = np.zeros(iters + 1, np.float32)
    lib().orc_mlem3d_pts_corr(ctypes.byref(im), nang, nr, _p(f32(pts)), len(za), _p(f32(za)),
                              _p(f32(zb)), _p(f32(meas)), _p(f32(init)), iters,
                              _p(f32(add)) if add is not None else None,
                              _p(f32(mult)) if mult is not None else None,
                              _p(img), _p(sens), _p(lik))
    shp = (im.nz, im.ny, im.nx)
    return img.reshape(shp), sens.reshape(shp), lik
