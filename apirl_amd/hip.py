"""ctypes binding of the HIP library (apirl_amd/libapirl_hip.so, C-ABI in
include/apirl_hip.h).

Device buffers are torch tensors on a ROCm device (torch is used only for device
memory and streams); the compute runs in the library's gfx950 kernels.  There is no
CPU fallback: if the library or a GPU is missing, every entry point raises."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapirl_hip.so")

APIRL_OK = 0
BP_SCATTER = 0
BP_GATHER = 1
BP_TILE = 2
RATIO_LIK_POSITIVE = 1
RATIO_ADDITIVE_LOCAL = 2
PREC_REFERENCE = 0
PREC_KAHAN = 1
PREC_F64 = 2          # all double
PREC_F64_ALPHA = 2    # earlier name of PREC_F64
PREC_F64_SETUP = 3    # double alpha setup, float stepping, Kahan float sum

_c_int = ctypes.c_int
_c_float = ctypes.c_float
_c_size = ctypes.c_size_t
_P = ctypes.c_void_p


class ImageDesc(ctypes.Structure):
    _fields_ = [("nx", _c_int), ("ny", _c_int), ("nz", _c_int), ("dx", _c_float),
                ("dy", _c_float), ("dz", _c_float)]


class SinoDesc(ctypes.Structure):
    _fields_ = [("dims", _c_int), ("nang", _c_int), ("nr", _c_int), ("nsino", _c_int),
                ("rfov_mm", _c_float), ("rscanner_mm", _c_float), ("ring_width_mm", _c_float),
                ("ang_deg", _P), ("r_mm", _P), ("z1_mean_mm", _P), ("z2_mean_mm", _P)]


class EmOptions(ctypes.Structure):
    _fields_ = [("iterations", _c_int), ("subsets", _c_int), ("bp_mode", _c_int),
                ("precision", _c_int), ("osem_rules", _c_int), ("additive_own_angles", _c_int),
                ("sensitivity_given", _c_int)]


class Tuning(ctypes.Structure):
    _fields_ = [("gather_zlanes", _c_int), ("gather_zvox", _c_int), ("gather_xcd", _c_int),
                ("tile_zs", _c_int),
                ("tile_cache_mb", _c_int)]


GATHER_XCD_ROUND_ROBIN = 1
GATHER_XCD_WEDGES = 2


class JobDesc(ctypes.Structure):
    _fields_ = [("plan", _P), ("iterations", _c_int), ("subsets", _c_int), ("osem_rules", _c_int),
                ("d_meas", _P), ("d_init", _P), ("d_additive", _P), ("d_mult", _P),
                ("d_out", _P), ("additive_own_angles", _c_int)]


# Every symbol the header declares, with its ctypes signature.
SIGNATURES = {
    "apirl_ctx_create": (_c_int, [_c_int, ctypes.POINTER(_P)]),
    "apirl_ctx_destroy": (_c_int, [_P]),
    "apirl_last_error": (ctypes.c_char_p, [_P]),
    "apirl_ctx_set_stream": (_c_int, [_P, _P]),
    "apirl_ctx_synchronize": (_c_int, [_P]),
    "apirl_malloc": (_c_int, [_P, ctypes.POINTER(_P), _c_size]),
    "apirl_free": (_c_int, [_P, _P]),
    "apirl_memcpy_h2d": (_c_int, [_P, _P, _P, _c_size]),
    "apirl_memcpy_d2h": (_c_int, [_P, _P, _P, _c_size]),
    "apirl_memset": (_c_int, [_P, _P, _c_int, _c_size]),
    "apirl_plan_create": (_c_int, [_P, ctypes.POINTER(ImageDesc), ctypes.POINTER(SinoDesc),
                                   ctypes.POINTER(_P)]),
    "apirl_plan_destroy": (_c_int, [_P]),
    "apirl_subset_num_angles": (_c_int, [_c_int, _c_int, _c_int]),
    "apirl_plan_subset_bins": (_c_size, [_P, _c_int, _c_int]),
    "apirl_plan_image_voxels": (_c_size, [_P]),
    "apirl_plan_intersections": (_c_int, [_P, _c_int, _c_int,
                                          ctypes.POINTER(ctypes.c_ulonglong)]),
    "apirl_plan_get_tuning": (_c_int, [_P, ctypes.POINTER(Tuning)]),
    "apirl_plan_set_tuning": (_c_int, [_P, ctypes.POINTER(Tuning)]),
    "apirl_project": (_c_int, [_P, _P, _P, _c_int, _c_int, _c_int]),
    "apirl_backproject": (_c_int, [_P, _P, _P, _c_int, _c_int, _c_int, _c_int]),
    "apirl_backproject_update": (_c_int, [_P, _P, _P, _P, _c_float, _c_int, _c_int, _c_int,
                                          _c_int, _c_int]),
    "apirl_divide_backproject": (_c_int, [_P, _P, _P, _P, _c_int, _c_int, _c_int]),
    "apirl_project_ratio": (_c_int, [_P, _P, _P, _P, _P, _P, _c_int, _c_int, _c_int, _c_int]),
    "apirl_project_ratio_noalloc": (_c_int, [_P, _P, _P, _P, _P, _P, _c_int, _c_int, _c_int,
                                             _c_int]),
    "apirl_em_update": (_c_int, [_P, _P, _P, _P, _c_size, _c_float, _c_int]),
    "apirl_sensitivity": (_c_int, [_P, _P, _P, _c_int, _c_int, _c_int, _c_int]),
    "apirl_mlem_threshold": (_c_int, [_P, _P, _c_size, ctypes.POINTER(_c_float)]),
    "apirl_em_reconstruct": (_c_int, [_P, ctypes.POINTER(EmOptions), _P, _P, _P, _P, _P, _P,
                                      _P]),
    "apirl_plan_set_oversampling": (_c_int, [_P, _c_int, _c_int]),
    "apirl_plan_set_scanner": (_c_int, [_P, _c_int]),
    "apirl_em_reconstruct_timed": (_c_int, [_P, ctypes.POINTER(EmOptions), _P, _P, _P, _P, _P,
                                            _P, _P, _P]),
    "apirl_batch_create": (_c_int, [_P, _c_int, _c_int, ctypes.POINTER(_P)]),
    "apirl_batch_destroy": (_c_int, [_P]),
    "apirl_batch_last_error": (ctypes.c_char_p, [_P]),
    "apirl_batch_job_bytes": (_c_size, [_P, _c_int]),
    "apirl_batch_job_backprojector": (_c_int, [_P]),
    "apirl_batch_set_launch": (_c_int, [_P, _c_int, _c_int]),
    "apirl_batch_set_schedule": (_c_int, [_P, _c_int, _c_int]),
    "apirl_batch_submit": (_c_int, [_P, ctypes.POINTER(JobDesc), ctypes.POINTER(_c_int)]),
    "apirl_batch_step": (_c_int, [_P, ctypes.POINTER(_c_int)]),
    "apirl_batch_job_done": (_c_int, [_P, _c_int, ctypes.POINTER(_c_int)]),
    "apirl_batch_synchronize": (_c_int, [_P]),
    "apirl_batch_release": (_c_int, [_P, _c_int]),
    "apirl_sino_grid": (None, [_c_int, _c_int, _c_float, _P, _P]),
    "apirl_ring_table": (_c_int, [_c_int, _c_float, _c_int, _P, _P, _P, _P, _P]),
}

_lib = None


class ApirlError(RuntimeError):
    pass


def load():
    """Loads libapirl_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ApirlError(f"{LIB_PATH} is missing: build it with `python -m apirl_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _np_ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def sino_grid(nang, nr, rfov):
    ang = np.zeros(nang, np.float32)
    r = np.zeros(nr, np.float32)
    load().apirl_sino_grid(nang, nr, rfov, _np_ptr(ang), _np_ptr(r))
    return ang, r


def segment_table(rings, span, mrd):
    """Segments ordered 0, +1, -1, +2, -2 ...: (sinograms per segment, min ring
    difference, max ring difference)."""
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
    """Mean ring coordinates (z1, z2) of every sinogram of a cylindrical 3D sinogram."""
    nsino, mn, mx = segment_table(rings, span, mrd)
    return ring_table_from(rings, zfov, nsino, mn, mx)


def ring_table_from(rings, zfov, nsino, mn, mx):
    """(z1, z2) per sinogram for an explicit segment table (as an Interfile header
    gives it: sinograms, minimum and maximum ring difference per segment)."""
    nsino = np.ascontiguousarray(nsino, np.int32)
    mn = np.ascontiguousarray(mn, np.int32)
    mx = np.ascontiguousarray(mx, np.int32)
    tot = int(nsino.sum())
    z1 = np.zeros(tot, np.float32)
    z2 = np.zeros(tot, np.float32)
    n = load().apirl_ring_table(rings, zfov, len(nsino), _np_ptr(nsino), _np_ptr(mn),
                                _np_ptr(mx), _np_ptr(z1), _np_ptr(z2))
    if n != tot:
        raise ApirlError("segment table does not match the ring enumeration")
    return z1, z2


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise ApirlError("no ROCm GPU visible: the HIP Siddon path needs an MI355X")
    return torch


class Context:
    """One device; work is enqueued on the null stream (torch's default stream)."""

    def __init__(self, device=0):
        _torch()
        self.lib = load()
        self.h = _P()
        st = self.lib.apirl_ctx_create(device, ctypes.byref(self.h))
        if st != APIRL_OK:
            raise ApirlError(f"apirl_ctx_create failed with status {st}")
        self.device = device

    def check(self, st, what):
        if st != APIRL_OK:
            msg = self.lib.apirl_last_error(self.h).decode()
            raise ApirlError(f"{what} failed (status {st}): {msg}")

    def synchronize(self):
        self.check(self.lib.apirl_ctx_synchronize(self.h), "synchronize")

    def close(self):
        if self.h:
            self.lib.apirl_ctx_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dptr(t):
    return _P(t.data_ptr()) if t is not None else None


class Plan:
    """Geometry plan: image grid + cylindrical sinogram (2D or 3D)."""

    def __init__(self, ctx, nx, ny, nz, dx, dz, nang, nr, rfov, rscanner, dims=3,
                 z1=None, z2=None, ring_width=0.0, ang=None, r=None):
        self.ctx = ctx
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nang, self.nr = nang, nr
        self.dims = dims
        self.nsino = 1 if dims == 2 else len(z1)
        im = ImageDesc(nx, ny, nz, dx, dx, dz)
        self._keep = []

        def keep(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.float32)
            self._keep.append(a)
            return _np_ptr(a)

        sd = SinoDesc(dims, nang, nr, self.nsino, rfov, rscanner, ring_width, keep(ang), keep(r),
                      keep(z1), keep(z2))
        self.h = _P()
        ctx.check(ctx.lib.apirl_plan_create(ctx.h, ctypes.byref(im), ctypes.byref(sd),
                                            ctypes.byref(self.h)), "apirl_plan_create")

    @classmethod
    def for_2d(cls, ctx, nx, dx, nang, nr, rfov, rscanner):
        return cls(ctx, nx, nx, 1, dx, 1.0, nang, nr, rfov, rscanner, dims=2)

    @classmethod
    def for_3d(cls, ctx, nx, ny, nz, dx, dz, nang, nr, rfov, rscanner, rings, zfov, span, mrd):
        z1, z2 = ring_table(rings, zfov, span, mrd)
        return cls(ctx, nx, ny, nz, dx, dz, nang, nr, rfov, rscanner, dims=3, z1=z1, z2=z2,
                   ring_width=np.float32(zfov) / np.float32(rings))

    def subset_angles(self, subset, nsubsets):
        return self.ctx.lib.apirl_subset_num_angles(self.nang, nsubsets, subset)

    def subset_shape(self, subset=0, nsubsets=1):
        return (self.nsino, self.subset_angles(subset, nsubsets), self.nr)

    def image_shape(self):
        return (self.nz, self.ny, self.nx)

    def intersections(self, subset=0, nsubsets=1):
        out = ctypes.c_ulonglong(0)
        self.ctx.check(self.ctx.lib.apirl_plan_intersections(self.h, subset, nsubsets,
                                                             ctypes.byref(out)),
                       "apirl_plan_intersections")
        return out.value

    @property
    def gather_zlanes(self):
        """3D gather voxel mapping: 1 = 16x16 slice tiles, Z = Z slices per wavefront."""
        t = Tuning()
        self.ctx.check(self.ctx.lib.apirl_plan_get_tuning(self.h, ctypes.byref(t)),
                       "apirl_plan_get_tuning")
        return t.gather_zlanes

    @gather_zlanes.setter
    def gather_zlanes(self, n):
        t = Tuning(gather_zlanes=int(n))
        self.ctx.check(self.ctx.lib.apirl_plan_set_tuning(self.h, ctypes.byref(t)),
                       "apirl_plan_set_tuning")

    SCANNERS = {"cylindrical": 0, "mmr": 1, "mmr_subset": 2}

    def set_scanner(self, name):
        """LOR geometry of a 3D plan: "cylindrical" (the plan's grids) or "mmr"
        (Siemens Biograph mMR: arc-corrected r, half-angle interleave, DOI radius)."""
        self.ctx.check(self.ctx.lib.apirl_plan_set_scanner(self.h, self.SCANNERS[name]),
                       "apirl_plan_set_scanner")

    def set_oversampling(self, ns, nas=1):
        """Trace every bin as ns x nas LORs (3D; the reference's SiddonProjector with
        numSamples / numAxialSamples), in the reference precision; the gather takes a
        bin's ns x nas LORs in turn, batches run such jobs with the single-job kernels."""
        self.ctx.check(self.ctx.lib.apirl_plan_set_oversampling(self.h, int(ns), int(nas)),
                       "apirl_plan_set_oversampling")

    def set_tuning(self, **fields):
        """Sets apirl_tuning fields (gather_zlanes, gather_zvox, gather_xcd, tile_zs,
        tile_cache_mb: -1 keeps no tile geometry); gather results do not depend on them."""
        t = Tuning(**{k: int(v) for k, v in fields.items()})
        self.ctx.check(self.ctx.lib.apirl_plan_set_tuning(self.h, ctypes.byref(t)),
                       "apirl_plan_set_tuning")

    def get_tuning(self):
        t = Tuning()
        self.ctx.check(self.ctx.lib.apirl_plan_get_tuning(self.h, ctypes.byref(t)),
                       "apirl_plan_get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    @property
    def gather_zvox(self):
        """3D gather voxels per lane along z (1..4)."""
        t = Tuning()
        self.ctx.check(self.ctx.lib.apirl_plan_get_tuning(self.h, ctypes.byref(t)),
                       "apirl_plan_get_tuning")
        return t.gather_zvox

    @gather_zvox.setter
    def gather_zvox(self, n):
        t = Tuning(gather_zvox=int(n))
        self.ctx.check(self.ctx.lib.apirl_plan_set_tuning(self.h, ctypes.byref(t)),
                       "apirl_plan_set_tuning")

    # --- Projector entry points on torch device tensors -----------------------
    def project(self, img, out=None, subset=0, nsubsets=1, precision=PREC_REFERENCE):
        torch = _torch()
        if out is None:
            out = torch.zeros(self.subset_shape(subset, nsubsets), dtype=torch.float32,
                              device=img.device)
        self._check_img(img)
        self._check_sino(out, subset, nsubsets)
        self.ctx.check(self.ctx.lib.apirl_project(self.h, _dptr(img), _dptr(out), subset,
                                                  nsubsets, precision), "apirl_project")
        return out

    def backproject(self, sino, out=None, subset=0, nsubsets=1, mode=BP_SCATTER,
                    precision=PREC_REFERENCE):
        torch = _torch()
        if out is None:
            out = torch.zeros(self.image_shape(), dtype=torch.float32, device=sino.device)
        self._check_img(out)
        self._check_sino(sino, subset, nsubsets)
        self.ctx.check(self.ctx.lib.apirl_backproject(self.h, _dptr(sino), _dptr(out), subset,
                                                      nsubsets, mode, precision),
                       "apirl_backproject")
        return out

    def backproject_update(self, sino, x, sens, thr=0.0, strict=0, subset=0, nsubsets=1,
                           precision=PREC_REFERENCE, mode=BP_GATHER):
        """x := cond ? x * (A^T sino) / sens : 0 in one gather or tile backprojection
        (apirl_backproject_update)."""
        self._check_img(x)
        self._check_img(sens)
        self._check_sino(sino, subset, nsubsets)
        self.ctx.check(self.ctx.lib.apirl_backproject_update(
            self.h, _dptr(sino), _dptr(x), _dptr(sens), float(thr), int(strict), subset,
            nsubsets, mode, precision), "apirl_backproject_update")
        return x

    def divide_backproject(self, meas, est, out=None, subset=0, nsubsets=1, mode=BP_SCATTER):
        torch = _torch()
        if out is None:
            out = torch.zeros(self.image_shape(), dtype=torch.float32, device=meas.device)
        self._check_img(out)
        self._check_sino(meas, subset, nsubsets)
        self._check_sino(est, subset, nsubsets)
        self.ctx.check(self.ctx.lib.apirl_divide_backproject(self.h, _dptr(meas), _dptr(est),
                                                             _dptr(out), subset, nsubsets, mode),
                       "apirl_divide_backproject")
        return out

    def project_ratio(self, img, meas, additive=None, subset=0, nsubsets=1, lik_positive_only=0,
                      precision=PREC_REFERENCE, ratio=None, want_lik=True, lik=None,
                      additive_local=False):
        """Fused forward step.  lik: optional float64 device tensor of one element that
        receives the likelihood (allocated here when want_lik and lik is None).
        additive_local: subset bin k takes the additive sinogram's angle k (the
        reference's OSEM, OsemSinogram3d.cpp:172-173) instead of its own angle."""
        torch = _torch()
        if ratio is None:
            ratio = torch.empty(self.subset_shape(subset, nsubsets), dtype=torch.float32,
                                device=img.device)
        if lik is not None:
            if lik.dtype != torch.float64 or lik.numel() != 1 or not lik.is_cuda:
                raise ApirlError("lik must be a float64 GPU tensor of one element")
        elif want_lik:
            lik = torch.zeros(1, dtype=torch.float64, device=img.device)
        self._check_img(img)
        self._check_sino(meas, 0, 1)
        self._check_sino(ratio, subset, nsubsets)
        if additive is not None:
            self._check_sino(additive, 0, 1)
        self.ctx.check(self.ctx.lib.apirl_project_ratio(
            self.h, _dptr(img), _dptr(meas), _dptr(additive), _dptr(ratio), _dptr(lik), subset,
            nsubsets, (RATIO_LIK_POSITIVE if lik_positive_only else 0) |
            (RATIO_ADDITIVE_LOCAL if additive_local else 0), precision), "apirl_project_ratio")
        return ratio, lik

    def sensitivity(self, out=None, mult=None, subset=0, nsubsets=1, mode=BP_SCATTER,
                    force_positive=1):
        torch = _torch()
        if out is None:
            out = torch.empty(self.image_shape(), dtype=torch.float32, device="cuda")
        self._check_img(out)
        self.ctx.check(self.ctx.lib.apirl_sensitivity(self.h, _dptr(mult), _dptr(out), subset,
                                                      nsubsets, mode, force_positive),
                       "apirl_sensitivity")
        return out

    def reconstruct(self, meas, init, iterations, subsets=1, bp_mode=BP_SCATTER,
                    precision=PREC_REFERENCE, osem_rules=None, additive=None, mult=None,
                    out=None, sens_out=None, times_out=None, additive_own_angles=False,
                    sens_in=None):
        """Device-resident MLEM (subsets=1) or OSEM; returns (image, likelihood[T*S+1]).
        times_out: optional float32 array [4, T*S] that receives each sub-iteration's
        time in ms: whole, projection + ratio, backprojection, update.
        additive_own_angles: add each subset bin's own angle of the additive sinogram
        instead of the reference's subset-local angle (equal for one subset).
        sens_in: sensitivity images [subsets, nz, ny, nx] to use instead of computing
        them (the reference's "sensitivity filename"); clamped in place under OSEM."""
        torch = _torch()
        if osem_rules is None:
            osem_rules = 1 if subsets > 1 else 0
        if out is None:
            out = torch.empty(self.image_shape(), dtype=torch.float32, device=init.device)
        if sens_in is not None:
            if sens_out is not None:
                raise ValueError("pass sens_in or sens_out, not both")
            if sens_in.numel() != subsets * self.nx * self.ny * self.nz or \
                    not sens_in.is_contiguous() or sens_in.dtype != torch.float32:
                raise ApirlError("sens_in must hold one float32 image per subset")
            sens_out = sens_in
        opt = EmOptions(iterations, subsets, bp_mode, precision, osem_rules,
                        1 if additive_own_angles else 0, 1 if sens_in is not None else 0)
        lik = np.zeros(iterations * subsets + 1, np.float64)
        self._check_img(init)
        self._check_img(out)
        self._check_sino(meas, 0, 1)
        if times_out is not None:
            if times_out.dtype != np.float32 or times_out.size != 4 * iterations * subsets or \
                    not times_out.flags.c_contiguous:
                raise ValueError("times_out must be a contiguous float32 array of 4*T*S")
            self.ctx.check(self.ctx.lib.apirl_em_reconstruct_timed(
                self.h, ctypes.byref(opt), _dptr(meas), _dptr(init), _dptr(additive),
                _dptr(mult), _dptr(out), _np_ptr(lik), _dptr(sens_out), _np_ptr(times_out)),
                "apirl_em_reconstruct_timed")
            return out, lik
        self.ctx.check(self.ctx.lib.apirl_em_reconstruct(
            self.h, ctypes.byref(opt), _dptr(meas), _dptr(init), _dptr(additive), _dptr(mult),
            _dptr(out), _np_ptr(lik), _dptr(sens_out)), "apirl_em_reconstruct")
        return out, lik

    def _check_img(self, t):
        if t.numel() != self.nx * self.ny * self.nz or not t.is_contiguous() or \
                t.dtype.__str__() != "torch.float32" or not t.is_cuda:
            raise ApirlError(f"image tensor must be contiguous float32 on the GPU with "
                             f"{self.nx * self.ny * self.nz} voxels")

    def _check_sino(self, t, subset, nsubsets):
        n = int(np.prod(self.subset_shape(subset, nsubsets)))
        if t.numel() != n or not t.is_contiguous() or t.dtype.__str__() != "torch.float32" or \
                not t.is_cuda:
            raise ApirlError(f"sinogram tensor must be contiguous float32 on the GPU with {n} bins")

    def close(self):
        if self.h:
            self.ctx.lib.apirl_plan_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Continuous batching of OSEM / MLEM jobs (apirl_batch_*): submit jobs at any
    time, call step() to run one sub-iteration of every active job, poll done()."""

    def __init__(self, ctx, max_jobs=64, max_sinograms=1024):
        self.ctx = ctx
        self.lib = ctx.lib
        self.h = _P()
        ctx.check(self.lib.apirl_batch_create(ctx.h, max_jobs, max_sinograms,
                                              ctypes.byref(self.h)), "apirl_batch_create")
        self._keep = {}

    def _check(self, st, what):
        if st != APIRL_OK:
            msg = self.lib.apirl_batch_last_error(self.h).decode()
            raise ApirlError(f"{what} failed (status {st}): {msg}")

    @staticmethod
    def job_bytes(plan, subsets):
        return plan.ctx.lib.apirl_batch_job_bytes(plan.h, subsets)

    @staticmethod
    def backprojector(plan):
        """The backprojector (BP_TILE or BP_GATHER) a batch runs a job of this plan with:
        its image equals plan.reconstruct(..., bp_mode=Batch.backprojector(plan))."""
        return plan.ctx.lib.apirl_batch_job_backprojector(plan.h)

    def set_launch(self, grid=0, use_graph=True):
        self._check(self.lib.apirl_batch_set_launch(self.h, grid, 1 if use_graph else 0),
                    "apirl_batch_set_launch")

    ORDER_SUBMISSION, ORDER_LARGEST_FIRST = 0, 1

    def set_schedule(self, grab=2, order=0):
        """Virtual blocks per work-counter draw (1..64) and the jobs' order in a round
        (ORDER_SUBMISSION or ORDER_LARGEST_FIRST); results do not depend on either."""
        self._check(self.lib.apirl_batch_set_schedule(self.h, int(grab), int(order)),
                    "apirl_batch_set_schedule")

    def submit(self, plan, meas, init, out, iterations, subsets=1, osem_rules=None,
               additive=None, mult=None, additive_own_angles=False):
        if osem_rules is None:
            osem_rules = 1 if subsets > 1 else 0
        plan._check_img(init)
        plan._check_img(out)
        plan._check_sino(meas, 0, 1)
        d = JobDesc(plan.h, iterations, subsets, osem_rules, _dptr(meas), _dptr(init),
                    _dptr(additive), _dptr(mult), _dptr(out), 1 if additive_own_angles else 0)
        jid = _c_int(-1)
        self._check(self.lib.apirl_batch_submit(self.h, ctypes.byref(d), ctypes.byref(jid)),
                    "apirl_batch_submit")
        self._keep[jid.value] = (plan, meas, init, out, additive, mult)
        return jid.value

    def step(self):
        n = _c_int(0)
        self._check(self.lib.apirl_batch_step(self.h, ctypes.byref(n)), "apirl_batch_step")
        return n.value

    def done(self, job_id):
        d = _c_int(0)
        self._check(self.lib.apirl_batch_job_done(self.h, job_id, ctypes.byref(d)),
                    "apirl_batch_job_done")
        return bool(d.value)

    def release(self, job_id):
        self._check(self.lib.apirl_batch_release(self.h, job_id), "apirl_batch_release")
        self._keep.pop(job_id, None)

    def synchronize(self):
        self._check(self.lib.apirl_batch_synchronize(self.h), "apirl_batch_synchronize")

    def run(self):
        """Steps until every submitted job has been scheduled, then waits."""
        while self.step() > 0:
            pass
        self.synchronize()

    def close(self):
        if self.h:
            self.lib.apirl_batch_destroy(self.h)
            self.h = _P()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def em_update(ctx, x, b, s, thr, strict):
    ctx.check(ctx.lib.apirl_em_update(ctx.h, _dptr(x), _dptr(b), _dptr(s), x.numel(), thr,
                                      int(strict)), "apirl_em_update")
    return x


def mlem_threshold(ctx, sens):
    out = ctypes.c_float(0)
    ctx.check(ctx.lib.apirl_mlem_threshold(ctx.h, _dptr(sens), sens.numel(), ctypes.byref(out)),
              "apirl_mlem_threshold")
    return out.value
