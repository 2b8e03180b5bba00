// Deterministic voxel-driven gather backprojector for gfx950.
//
// One voxel per lane.  For each angle of the subset the lane finds the contiguous
// radial bins whose LOR can cross its voxel column, and for 3D, per segment, the
// contiguous range of sinograms whose ray can cross its slice.  For every candidate
// ray it reconstructs, without walking the ray, the state the reference's Siddon
// loop would be in when it emits this voxel (recon/src/Siddon.cpp:310-340, 615-639):
//   (a, b, c) = number of x, y, z crossings taken before the voxel,
//   the last crossing taken (alpha_c) and the next one (the segment end),
// using the exact values of the float accumulators (alpha_exact.h).  The voxel is
// on the ray's list iff every crossing taken precedes every crossing not taken in
// the loop's order (value, then x before y before z) and a + b + c is below the
// list length |di| + |dj| + |dk| (:220, :521), which reproduces the dropped exit
// voxel and the zero-length entry segments.  The weight (next - alpha_c) * L is then
// the reference's float, bit for bit.
//
// Terms are summed in a fixed order (angle, radial bin, segment, sinogram), so the
// result does not depend on scheduling: repeated runs are bit-identical.  No atomics.
#include <cfloat>
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "apirl_internal.h"
#include "alpha_exact.h"
#include "siddon_f64.h"

namespace apirl {

namespace {

constexpr int kGBlock = 256;
// axial margin of the sinogram search: covers the drift of the accumulated z
// crossings against the exact slab faces (well below 1e-3 mm here)
constexpr float kZMargin = 0.05f;

__device__ __forceinline__ float z_at(float za, float vz, float alpha) { return za + vz * alpha; }

// ZRay (apirl_internal.h): per-ray axial part of the Siddon setup for one subset
// (Siddon.cpp:150-182, 220, 244-255, 279-286) and the bin value, computed once per
// backprojection by zsetup_kernel.

// The axial part of one ray's setup in float (the reference's operations in its
// order): entry and exit slices, first z crossing and its increment, and the length.
__device__ __forceinline__ void zray_setup(const Geo& g, const RayXY& R, float za, float zb,
                                           ZRay& z) {
  const float vz = zb - za;
  const float z1 = za + vz * R.amin;
  const float z2 = za + vz * R.amax;
  const int kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
  z.kinc = (short)kinc;
  z.k0 = (int)floorf((z1 - g.offz) / g.dz);
  z.k1 = (short)(int)floorf((z2 - g.offz) / g.dz);
  z.azu = fabsf(g.dz / vz);
  if (vz > 0) z.az0 = (g.offz + (z.k0 + kinc) * g.dz - za) / vz;
  else if (vz < 0) z.az0 = (g.offz + z.k0 * g.dz - za) / vz;
  else z.az0 = FLT_MAX;
  const float ez = (za + vz) - za;
  z.L = sqrtf(R.lxy + ez * ez);
}

// nz_pair[ka * nr + r] != 0 iff some sinogram has a non-zero bin at (angle ka, r):
// the gather skips (angle, r) pairs whose bins are all zero (zero bins add nothing,
// SiddonProjector.cpp:496).
__device__ __forceinline__ void zsetup_one(const Geo& g, const RayXY* __restrict__ rays,
                                           const float2* __restrict__ zab,
                                           const float* __restrict__ sino, int kind,
                                           int nang_sub, int subset, int nsubsets,
                                           ZRay* __restrict__ out, int* __restrict__ nz_pair,
                                           float* __restrict__ ztab,
                                           const int* __restrict__ zsoff, int zblock,
                                           long idx, int f64s = 0) {
  const int rr = (int)(idx % g.nr);
  const long t = idx / g.nr;
  const int ka = (int)(t % nang_sub);
  const int s = (int)(t / nang_sub);
  const int a = subset + nsubsets * ka;
  const RayXY R = rays[(long)a * g.nr + rr];
  const float za = zab[s].x, zb = zab[s].y;
  const float vz = zb - za;
  ZRay z;
  const int kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
  z.kinc = (short)kinc;
  if (f64s) {
    // APIRL_PREC_F64_SETUP: the axial setup in double from the float LOR, rounded to
    // float, as the projector's F64_SETUP loop starts from it
    SetupXY64 xy;
    setup_xy_f64(g, R.p0x, R.p0y, R.vx, R.vy, xy);
    SetupZ64 zs;
    setup_z_f64(g, xy, za, R.vx, R.vy, (double)vz, zs);
    z.k0 = zs.k0;
    z.k1 = (short)zs.k1;
    z.az0 = (float)zs.az;
    z.azu = (float)zs.azu;
    z.L = (float)zs.L;
  } else {
    zray_setup(g, R, za, zb, z);
  }
  float b = 1.f;
  if (kind == BP_PLAIN) b = sino[idx];
  else if (kind == BP_MULT) b = sino[((long)s * g.nang + a) * g.nr + rr];
  z.b = b;
  // The ray's z crossings as the reference accumulates them (alpha_z += alpha_z_u,
  // Siddon.cpp:338): entry t is the value after t additions.  A ray with more
  // crossings than its slot (not expected: the slot bounds |Vz| (amax - amin) / dz)
  // is flagged and the gather reconstructs its crossings with alpha_after.
  // Entries are written for the slices k0 + kinc, k0 + 2 kinc, ... as far as the
  // slot reaches, at most |k1 - k0| + 2 of them (the loop's N steps can include a z
  // step more than the floor-based count when an x or y step fewer is taken);
  // zcount records how many, and the gather reconstructs any crossing beyond them
  // with alpha_after.
  const int nzc = abs((int)z.k1 - z.k0);
  int nw = 0;
  // a zero bin adds nothing (the gathers skip it once they have its record), so its
  // crossings are not written
  if (ztab && kinc != 0 && b != 0.f) {
    const int base = zsoff[s], slot = zsoff[s + 1] - base, kmin = zsoff[g.nsino + 1 + s];
    // entry e of the sinogram's slot lives at ((base + e) * nang_sub + ka) * nr + rr:
    // the rays of one (sinogram, angle) are adjacent, as the gather's lanes read them
    const long npair = (long)nang_sub * g.nr;
    float* dst = ztab + (long)ka * g.nr + rr;
    float az = z.az0;
    int kk = z.k0 + kinc;
    while (nw < nzc + 2 && kk - kmin >= 0 && kk - kmin < slot) {
      dst[(long)(base + kk - kmin) * npair] = az;  // entering slice kk: value after nw additions
      az = az + z.azu;
      kk += kinc;
      nw++;
    }
  }
  z.zcount = nw;
  z.pad = 0;
  out[idx] = z;  // (sinogram, angle, radial bin) order
  if (b != 0.f && (R.flags & RAY_VALID)) {
    int* f = &nz_pair[(long)ka * g.nr + rr];
    if (*f == 0) atomicOr(f, 1);
  }
}

__global__ __launch_bounds__(kGBlock) void zsetup_kernel(Geo g, const RayXY* __restrict__ rays,
                                                         const float2* __restrict__ zab,
                                                         const float* __restrict__ sino, int kind,
                                                         int nang_sub, int subset, int nsubsets,
                                                         long nrays, ZRay* __restrict__ out,
                                                         int* __restrict__ nz_pair,
                                                         float* __restrict__ ztab,
                                                         const int* __restrict__ zsoff,
                                                         int zblock, int f64s) {
  const long idx = (long)blockIdx.x * kGBlock + threadIdx.x;
  if (idx < nrays)
    zsetup_one(g, rays, zab, sino, kind, nang_sub, subset, nsubsets, out, nz_pair, ztab, zsoff,
               zblock, idx, f64s);
}

__global__ __launch_bounds__(kGBlock) void clear_int_kernel(int* __restrict__ p, long n) {
  const long i = (long)blockIdx.x * kGBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}



// Voxel tile of one workgroup in the transaxial mapping: 16 x 16 voxels of one slice.
constexpr int kTile = 16;
constexpr int kWave = 64;

// Per-voxel accumulator of the gather for each precision mode.
template <int PREC>
struct Acc {
  using alpha_t = float;
  float s;
  __device__ void init(float v) { s = v; }
  __device__ void add(float t) { s = s + t; }
  // add(t) when c, else nothing (selected, so no branch)
  __device__ void add_if(bool c, float t) {
    const float n = s + t;
    s = c ? n : s;
  }
  __device__ void add_f64(double) {}
  __device__ float value() const { return s; }
};
template <>
struct Acc<APIRL_PREC_KAHAN> {
  using alpha_t = float;
  float s, c;
  __device__ void init(float v) { s = v; c = 0.f; }
  __device__ void add(float t) {
    const float y = t - c;
    const float u = s + y;
    c = (u - s) - y;
    s = u;
  }
  __device__ void add_if(bool cond, float t) {
    const float y = t - c;
    const float u = s + y;
    const float nc = (u - s) - y;
    s = cond ? u : s;
    c = cond ? nc : c;
  }
  __device__ void add_f64(double) {}
  __device__ float value() const { return s; }
};
// F64_SETUP: float crossings and weights from the double setup, summed as KAHAN.
template <>
struct Acc<APIRL_PREC_F64_SETUP> : Acc<APIRL_PREC_KAHAN> {};
template <>
struct Acc<APIRL_PREC_F64_ALPHA> {
  using alpha_t = double;
  double s;
  __device__ void init(float v) { s = v; }
  __device__ void add(float t) { s += t; }
  __device__ void add_f64(double t) { s += t; }
  __device__ float value() const { return (float)s; }
};

// First x / y crossings and increments of a ray.  REFERENCE and KAHAN use the
// reference's float values from the ray table (Siddon.cpp:253-286, with the 2D
// clamp of a negative first crossing to alpha_min, :590-606); F64_ALPHA in 2D
// recomputes them in double from the float entry voxel (3D F64_ALPHA does the whole
// setup in double, siddon_f64.h).
template <int PREC, int DIMS>
__device__ __forceinline__ void first_crossings(const Geo& g, const RayXY& t, int iinc, int jinc,
                                                typename Acc<PREC>::alpha_t& ax0,
                                                typename Acc<PREC>::alpha_t& axu,
                                                typename Acc<PREC>::alpha_t& ay0,
                                                typename Acc<PREC>::alpha_t& ayu) {
  if (PREC != APIRL_PREC_F64_ALPHA) {
    ax0 = t.ax;
    axu = t.axu;
    ay0 = t.ay;
    ayu = t.ayu;
  } else {
    const double R = g.R;
    const double vx = t.vx, vy = t.vy;
    ax0 = FLT_MAX;
    ay0 = FLT_MAX;
    axu = 0.0;
    ayu = 0.0;
    if (iinc != 0) {
      axu = fabs((double)g.dx / vx);
      ax0 = (-R + (double)(t.i0 + (iinc > 0 ? 1 : 0)) * g.dx - t.p0x) / vx;
      if (DIMS == 2 && ax0 < 0.0) ax0 = t.amin;
    }
    if (jinc != 0) {
      ayu = fabs((double)g.dy / vy);
      ay0 = (-R + (double)(t.j0 + (jinc > 0 ? 1 : 0)) * g.dy - t.p0y) / vy;
      if (DIMS == 2 && ay0 < 0.0) ay0 = t.amin;
    }
  }
}

// n-th crossing after the first: the exact float accumulator (alpha_exact.h) or,
// in double, a0 + n u.
// alpha_after out of line: the column kernel needs it only for the rare crossings the
// table does not hold, and inlined, parts of it are hoisted to every sinogram.
__device__ __noinline__ float alpha_after_call(float a, float u, int n) {
  return alpha_after(a, u, n);
}

template <int PREC>
__device__ __forceinline__ typename Acc<PREC>::alpha_t crossing(typename Acc<PREC>::alpha_t a0,
                                                               typename Acc<PREC>::alpha_t u,
                                                               int n) {
  if (PREC == APIRL_PREC_F64_ALPHA) return a0 + (double)n * u;
  else return alpha_after_call((float)a0, (float)u, n);
}

template <int PREC>
__device__ __forceinline__ void z_crossings(const Geo& g, const ZRay& z, float2 zz,
                                            typename Acc<PREC>::alpha_t& az0,
                                            typename Acc<PREC>::alpha_t& azu) {
  if (PREC != APIRL_PREC_F64_ALPHA) {
    az0 = z.az0;
    azu = z.azu;
  } else {
    const double za = zz.x, vz = (double)zz.y - zz.x;
    azu = fabs((double)g.dz / vz);
    if (vz > 0) az0 = ((double)g.offz + (double)(z.k0 + 1) * g.dz - za) / vz;
    else if (vz < 0) az0 = ((double)g.offz + (double)z.k0 * g.dz - za) / vz;
    else az0 = FLT_MAX;
  }
}

// Range [jlo, jhi] of bins whose LORs can pass within hw of a point at signed
// distance sc from the centre along their normal, from the radial lookup table
// (one grid step of slack on each side for rounding); empty when jlo > jhi.
__device__ __forceinline__ void lut_bins(const GatherArgs& A, float sc, float hw, int& jlo,
                                         int& jhi) {
  int kl = (int)floorf((sc - hw - A.rl0) * A.rl_inv_h) - 1;
  int kh = (int)ceilf((sc + hw - A.rl0) * A.rl_inv_h) + 1;
  kl = kl < 0 ? 0 : (kl > A.nrl - 1 ? A.nrl - 1 : kl);
  kh = kh < 0 ? 0 : (kh > A.nrl - 1 ? A.nrl - 1 : kh);
  jlo = A.rlut[kl].x;
  jhi = A.rlut[kh].y;
}

// Radial grid of a gather launch: known at compile time for single-job launches
// (uniform, or the lookup table), read from the arguments otherwise.
enum GatherGrid { kGridAny = 0, kGridUniform = 1, kGridLut = 2 };

template <int GRID>
__device__ __forceinline__ bool grid_lut(const GatherArgs& A) {
  return GRID == kGridLut || (GRID == kGridAny && A.rlut != nullptr);
}

// Candidate bins of angle a for a voxel column centred at (xc, yc): the bins whose
// LOR passes within the column's half-width (plus the float margin) of its centre.
// Even bins (lo0, hi0) and odd bins (lo1, hi1) get their own range; on the plan's
// uniform grid, where every bin of an angle shares it, the two ranges are equal.
template <int GRID>
__device__ __forceinline__ void candidate_bins(const GatherArgs& A, int a, float xc, float yc,
                                               float inv_dr, int& lo0, int& hi0, int& lo1,
                                               int& hi1) {
  const Geo& g = A.g;
  const float2 cs = A.angcs[a];
  const float sc = xc * cs.x + yc * cs.y;
  const float hw = 0.5f * g.dx * (fabsf(cs.x) + fabsf(cs.y)) + g.rmargin;
  if (!grid_lut<GRID>(A)) {
    int jlo = (int)ceilf((sc - hw - g.r0) * inv_dr);
    int jhi = (int)floorf((sc + hw - g.r0) * inv_dr);
    lo0 = lo1 = jlo < 0 ? 0 : jlo;
    hi0 = hi1 = jhi > g.nr - 1 ? g.nr - 1 : jhi;
    return;
  }
  lut_bins(A, sc, hw, lo0, hi0);
  const float2 co = A.angcs_odd[a];
  const float so = xc * co.x + yc * co.y;
  const float hwo = 0.5f * g.dx * (fabsf(co.x) + fabsf(co.y)) + g.rmargin;
  lut_bins(A, so, hwo, lo1, hi1);
}

// Bin rr is a candidate of its parity's range.
__device__ __forceinline__ bool in_parity_range(int rr, int lo0, int hi0, int lo1, int hi1) {
  return (rr & 1) ? (rr >= lo1 && rr <= hi1) : (rr >= lo0 && rr <= hi0);
}

// A voxel's backprojection b: stored, or with the fused EM update
// (MlemSinogram3d.cpp:228-242, OsemSinogram3d.cpp:217-237, Mlem2d.cpp:171-185) applied
// to the image voxel it belongs to, as em_update_kernel computes it.
__device__ __forceinline__ void store_voxel(const GatherArgs& A, long v, float b) {
  if (!A.upd_sens) {
    A.img[v] = b;
    return;
  }
  const float sv = A.upd_sens[v];
  const bool ok = A.upd_strict ? (sv > A.upd_thr) : (sv >= A.upd_thr);
  A.img[v] = ok ? A.img[v] * b / sv : 0.f;
}

// Gather of one voxel (i, j, k): adds its terms to A.img[v] in the fixed order
// (angle, radial bin, segment, sinogram).  s_zab is the axial table in LDS (3D).
template <int DIMS, int PREC>
__device__ __forceinline__ void gather_voxel(const GatherArgs& A, const float2* s_zab, int i,
                                             int j, int k) {
  const Geo& g = A.g;
  const int* s_soff = (const int*)(s_zab + g.nsino);  // slot offsets follow the axial table
  const int* s_kmin = s_soff + g.nsino + 1;           // then the slots' first slices
  const long v = ((long)k * g.ny + j) * g.nx + i;
  // voxel centre for the candidate search
  const float xc = -g.R + g.dx * ((float)i + 0.5f);
  const float yc = -g.R + g.dy * ((float)j + 0.5f);
  const float zk0 = g.offz + k * g.dz, zk1 = zk0 + g.dz;
  const float inv_dr = 1.0f / g.dr;
  // Accumulator: float in the reference order (REFERENCE), float with a Kahan
  // compensation term (KAHAN), or double (F64_ALPHA).
  Acc<PREC> acc;
  acc.init(A.upd_sens ? 0.f : A.img[v]);
  for (int ka = 0; ka < A.nang_sub; ka++) {
    const int a = A.subset + A.nsubsets * ka;
    int lo0, hi0, lo1, hi1;
    candidate_bins<kGridAny>(A, a, xc, yc, inv_dr, lo0, hi0, lo1, hi1);
    const int jlo = lo0 < lo1 ? lo0 : lo1, jhi = hi0 > hi1 ? hi0 : hi1;
    for (int rr = jlo; rr <= jhi; rr++) {
      if (A.rlut && !in_parity_range(rr, lo0, hi0, lo1, hi1)) continue;
      if (DIMS == 3 && A.nz_pair[(long)ka * g.nr + rr] == 0) continue;
      const RayXY t = A.rays[(long)a * g.nr + rr];
      if (!(t.flags & RAY_VALID)) continue;
      const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
      // Entry voxel, transaxial list length and alpha range of the ray: the
      // reference's float setup from the ray table, or (F64_ALPHA, 3D) the same
      // setup carried out in double (siddon_f64.h).
      using R = typename Acc<PREC>::alpha_t;
      constexpr bool kF64Setup =
          (PREC == APIRL_PREC_F64_ALPHA || PREC == APIRL_PREC_F64_SETUP) && DIMS == 3;
      // crossings from the (double) setup by repeated addition rather than the plan's
      // float-setup tables
      constexpr bool kOwnCrossings = PREC == APIRL_PREC_F64_ALPHA || kF64Setup;
      int ti0 = t.i0, tj0 = t.j0, tnxy = t.nxy;
      R tamin = t.amin, tamax = t.amax;
      R ax0, axu, ay0, ayu;
      SetupXY64 xy;
      if (kF64Setup) {
        setup_xy_f64(g, t.p0x, t.p0y, t.vx, t.vy, xy);
        ti0 = xy.i0;
        tj0 = xy.j0;
        tnxy = abs(xy.i1 - xy.i0) + abs(xy.j1 - xy.j0);
        tamin = (R)xy.amin;
        tamax = (R)xy.amax;
        ax0 = (R)xy.ax;
        axu = (R)xy.axu;
        ay0 = (R)xy.ay;
        ayu = (R)xy.ayu;
      } else {
        first_crossings<PREC, DIMS>(g, t, iinc, jinc, ax0, axu, ay0, ayu);
      }
      int na, nb;
      if (iinc == 0) {
        if (i != ti0) continue;
        na = 0;
      } else {
        na = (i - ti0) * iinc;
        if (na < 0) continue;
      }
      if (jinc == 0) {
        if (j != tj0) continue;
        nb = 0;
      } else {
        nb = (j - tj0) * jinc;
        if (nb < 0) continue;
      }
      if (na + nb >= tnxy + (DIMS == 3 ? g.nz + 1 : 0)) continue;
      // x / y crossings around the column: the float accumulators of the reference
      // loop reconstructed exactly, or (F64_ALPHA) the crossings in double
      R Xp = -FLT_MAX, Xn, Yp = -FLT_MAX, Yn;
      Xn = ax0;
      Yn = ay0;
      if (kOwnCrossings) {
        if (na > 0) {
          Xp = crossing<PREC>(ax0, axu, na - 1);
          Xn = Xp + axu;
        }
        if (nb > 0) {
          Yp = crossing<PREC>(ay0, ayu, nb - 1);
          Yn = Yp + ayu;
        }
      } else {
        // the accumulated crossings by which the ray enters column i and row j
        // (plan-level table, written by repeated float addition)
        const float* xy = A.xytab + ((long)a * g.nr + rr) * (g.nx + g.ny);
        if (na > 0) {
          Xp = xy[i];
          Xn = Xp + axu;
        }
        if (nb > 0) {
          Yp = xy[g.nx + j];
          Yn = Yp + ayu;
        }
      }
      // taken x before untaken y (x wins ties), taken y before untaken x (strict)
      if (na > 0 && !(Xp <= Yn)) continue;
      if (nb > 0 && !(Yp < Xn)) continue;
      if (DIMS == 2) {
        const int m = na + nb;
        if (m >= tnxy) continue;
        R ac = tamin;
        if (m > 0) ac = Xp > Yp ? Xp : Yp;
        const R an = Xn < Yn ? Xn : Yn;
        float b = 1.f;
        if (A.kind == BP_PLAIN) b = A.sino[(long)ka * g.nr + rr];
        else if (A.kind == BP_MULT) b = A.sino[(long)a * g.nr + rr];
        if (PREC == APIRL_PREC_F64_ALPHA) {
          acc.add_f64((an - ac) * sqrt((double)t.vx * t.vx + (double)t.vy * t.vy) * b);
        } else {
          const float w = (float)(an - ac) * sqrtf(t.lxy);
          acc.add(w * b * 1.0f);  // SiddonProjector.cpp:72
        }
        continue;
      }
      // alpha interval of the column; the sinograms whose ray crosses the slab
      // [zk0, zk1] inside it form a contiguous run of each segment (the mean ring
      // coordinates grow with the index), found by a short scan from an estimate.
      const float lo = fmaxf((float)tamin, fmaxf((float)Xp, (float)Yp));
      const float hi = fminf((float)tamax, fminf((float)Xn, (float)Yn));
      const float amid = 0.5f * (lo + hi) - 0.5f;
      const float zlo_t = zk0 - kZMargin, zhi_t = zk1 + kZMargin;
      for (int sg = 0; sg < A.nseg; sg++) {
        const GatherSeg S = A.segs[sg];
        int sl = (int)floorf((0.5f * (zk0 + zk1) - amid * S.dlo - S.z0) * S.dhi + 0.5f);
        sl = sl < 0 ? 0 : (sl > S.count - 1 ? S.count - 1 : sl);
        // first sinogram whose ray reaches the slab's lower face
        while (sl > 0) {
          const float2 z = s_zab[S.first + sl - 1];
          const float vz = z.y - z.x;
          if (fmaxf(z.x + vz * lo, z.x + vz * hi) < zlo_t) break;
          sl--;
        }
        for (; sl < S.count; sl++) {
          const int s = S.first + sl;
          const float2 zz = s_zab[s];
          const float vzz = zz.y - zz.x;
          const float za_lo = z_at(zz.x, vzz, lo), za_hi = z_at(zz.x, vzz, hi);
          if (fmaxf(za_lo, za_hi) < zlo_t) continue;
          if (fminf(za_lo, za_hi) > zhi_t) break;
          const ZRay z = A.zr[((long)s * A.nang_sub + ka) * g.nr + rr];
          // the table entry for this slice, loaded alongside the record (its address
          // does not depend on it); used only when the ray enters the slice by a z
          // crossing of its own, which makes it the entry zsetup wrote
          float ztv = 0.f;
          if (!kOwnCrossings && A.ztab) {
            const int slot = s_soff[s + 1] - s_soff[s];
            int kk = k - s_kmin[s];
            kk = kk < 0 ? 0 : (kk >= slot ? slot - 1 : kk);
            ztv = A.ztab[((long)(s_soff[s] + kk) * A.nang_sub + ka) * g.nr + rr];
          }
          if (z.b == 0.f) continue;  // SiddonProjector.cpp:496
          int zk0 = z.k0, zk1 = z.k1, zkinc = z.kinc;
          R az0, azu;
          double Ld = 0.0;
          float Lf = z.L;
          if (kF64Setup) {
            SetupZ64 zs;
            setup_z_f64(g, xy, zz.x, t.vx, t.vy, (double)(zz.y - zz.x), zs);
            zk0 = zs.k0;
            zk1 = zs.k1;
            zkinc = zs.kinc;
            az0 = (R)zs.az;
            azu = (R)zs.azu;
            Ld = zs.L;
            Lf = (float)zs.L;
          } else {
            z_crossings<PREC>(g, z, zz, az0, azu);
          }
          int nc;
          if (zkinc == 0) {
            if (k != zk0) continue;
            nc = 0;
          } else {
            nc = (k - zk0) * zkinc;
            if (nc < 0) continue;
          }
          const int m = na + nb + nc;
          if (m >= tnxy + abs(zk1 - zk0)) continue;
          R Zp = -FLT_MAX, Zn = az0;
          if (nc > 0) {
            if (kOwnCrossings) {
              Zp = crossing<PREC>(az0, azu, nc - 1);
            } else {
              Zp = nc <= z.zcount ? ztv : crossing<PREC>(az0, azu, nc - 1);
            }
            Zn = Zp + azu;
          }
          // cross-axis order (x before y before z on ties)
          if (na > 0 && !(Xp <= Zn)) continue;
          if (nb > 0 && !(Yp <= Zn)) continue;
          if (nc > 0 && !(Zp < Xn && Zp < Yn)) continue;
          R ac = tamin;
          if (m > 0) {
            ac = Xp > Yp ? Xp : Yp;
            ac = ac > Zp ? ac : Zp;
          }
          R an = Xn < Yn ? Xn : Yn;
          an = an < Zn ? an : Zn;
          if (PREC == APIRL_PREC_F64_ALPHA) {
            acc.add_f64((an - ac) * Ld * z.b);
          } else {
            const float w = (float)(an - ac) * Lf;
            acc.add(w * 1.0f * z.b);  // SiddonProjector.cpp:546
          }
        }
      }
    }
  }
  store_voxel(A, v, acc.value());
}

// One 32-byte record as two 16-byte loads.
__device__ __forceinline__ ZRay load_zray(const ZRay* p) {
  const int4* q = reinterpret_cast<const int4*>(p);
  const int4 a = q[0], b = q[1];
  ZRay z;
  z.k0 = a.x;
  z.k1 = (short)(a.y & 0xffff);
  z.kinc = (short)(a.y >> 16);
  z.azu = __int_as_float(a.z);
  z.az0 = __int_as_float(a.w);
  z.L = __int_as_float(b.x);
  z.b = __int_as_float(b.y);
  z.zcount = b.z;
  z.pad = b.w;
  return z;
}

// LDS of the column kernel: the axial table (float2 per sinogram), per sinogram where
// slice k's crossing-table entry lives (entry index clamp(k + x, y, z) with x = slot
// start - first slice, [y, z] = the slot's entries; int4, 16-byte aligned), and the
// segments' sinogram runs (read once per candidate ray and segment: from LDS rather
// than by scalar loads, 2.53 -> 2.48 ms per C2 subset).
__host__ __device__ inline size_t column_lds_off(int nsino) { return ((size_t)nsino * sizeof(float2) + 15) / 16 * 16; }
inline size_t column_lds_bytes(int nsino, int nseg = 0) {
  return column_lds_off(nsino) + (size_t)nsino * sizeof(int4) + (size_t)nseg * sizeof(GatherSeg);
}

// 3D gather of ZV voxels of one column, (i, j, kb) .. (i, j, kb + ZV - 1), by one lane
// (float precision modes).  Each voxel receives exactly the terms gather_voxel gives
// it, in the same order (angle, radial bin, segment, sinogram): the voxels share the
// transaxial part of every candidate ray (its column tests and x / y crossings) and
// one scan of each segment's sinograms over the union of their slabs.  Inside the
// scan, each sinogram is tested against each voxel's own slab exactly as
// gather_voxel tests it.  Record and table addresses are formed from a per-(angle,
// radial bin) base and 32-bit offsets s npair and e npair, formed with 24-bit
// multiplies: gather_fits admits a subset only when its records number below 2^31,
// the z-table offsets stay below 2^32 (the table is held under 1 GiB) and
// npair, the sinograms and the table slots stay below 2^24, which makes both exact.
template <int PREC, int ZV, int GRID = kGridAny>
__device__ __forceinline__ void gather_column(const GatherArgs& A, const float2* s_zab,
                                              const int4* s_slot, const GatherSeg* segs, int i,
                                              int j, int kb) {
  static_assert(PREC != APIRL_PREC_F64_ALPHA && PREC != APIRL_PREC_F64_SETUP,
                "double-precision setup uses gather_voxel");
  const Geo& g = A.g;
  const int nk = g.nz - kb < ZV ? g.nz - kb : ZV;  // voxels of this column segment (>= 1)
  const long nxy = (long)g.nx * g.ny;
  const long v0 = ((long)kb * g.ny + j) * g.nx + i;
  const float xc = -g.R + g.dx * ((float)i + 0.5f);
  const float yc = -g.R + g.dy * ((float)j + 0.5f);
  const float inv_dr = 1.0f / g.dr;
  const unsigned npair = (unsigned)A.nang_sub * (unsigned)g.nr;
  // each voxel's slab and search bounds, as gather_voxel computes them
  float zlo_t[ZV], zhi_t[ZV];
  Acc<PREC> acc[ZV];
#pragma unroll
  for (int q = 0; q < ZV; q++) {
    const int k = kb + q;
    const float zk0 = g.offz + k * g.dz, zk1 = zk0 + g.dz;
    zlo_t[q] = zk0 - kZMargin;
    zhi_t[q] = zk1 + kZMargin;
    acc[q].init(q < nk && !A.upd_sens ? A.img[v0 + q * nxy] : 0.f);
  }
  const float zlo_all = zlo_t[0];
  float zhi_all = zhi_t[0];
#pragma unroll
  for (int q = 1; q < ZV; q++)
    if (q < nk) zhi_all = zhi_t[q];
  const float zmid_all = 0.5f * (zlo_all + zhi_all);
  for (int ka = 0; ka < A.nang_sub; ka++) {
    const int a = A.subset + A.nsubsets * ka;
    int lo0, hi0, lo1, hi1;
    candidate_bins<GRID>(A, a, xc, yc, inv_dr, lo0, hi0, lo1, hi1);
    const int jlo = lo0 < lo1 ? lo0 : lo1, jhi = hi0 > hi1 ? hi0 : hi1;
    for (int rr = jlo; rr <= jhi; rr++) {
      if (grid_lut<GRID>(A) && !in_parity_range(rr, lo0, hi0, lo1, hi1)) continue;
      const unsigned pair = (unsigned)ka * (unsigned)g.nr + (unsigned)rr;
      if (A.nz_pair[pair] == 0) continue;
      const RayXY t = A.rays[(long)a * g.nr + rr];
      const float* xy = A.xytab + ((long)a * g.nr + rr) * (g.nx + g.ny);
      // entry crossings of column i and row j (table entries exist for every i, j, so
      // they are loaded before the tests that decide whether they are used)
      const float xyi = xy[i], xyj = xy[g.nx + j];
      const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
      const int tnxy = t.nxy;
      const float tamin = t.amin, tamax = t.amax;
      float ax0, axu, ay0, ayu;
      first_crossings<PREC, 3>(g, t, iinc, jinc, ax0, axu, ay0, ayu);
      // the transaxial tests of gather_voxel as one predicate
      // crossings already taken in x and y (increments are -1, 0 or 1: a 24-bit
      // multiply is exact and full rate)
      const int na = __mul24(i - t.i0, iinc), nb = __mul24(j - t.j0, jinc);
      // (iinc == 0 ? i == i0 : na >= 0) and the same for j, as lane-mask operations
      bool okt = (t.flags & RAY_VALID) != 0;
      okt &= (i == t.i0) | ((iinc != 0) & (na >= 0));
      okt &= (j == t.j0) | ((jinc != 0) & (nb >= 0));
      okt &= na + nb < tnxy + g.nz + 1;
      const float Xp = na > 0 ? xyi : -FLT_MAX, Xn = na > 0 ? xyi + axu : ax0;
      const float Yp = nb > 0 ? xyj : -FLT_MAX, Yn = nb > 0 ? xyj + ayu : ay0;
      okt &= ((na <= 0) | (Xp <= Yn)) & ((nb <= 0) | (Yp < Xn));
      if (!okt) continue;
      const float lo = fmaxf(tamin, fmaxf(Xp, Yp));
      const float hi = fminf(tamax, fminf(Xn, Yn));
      const float amid = 0.5f * (lo + hi) - 0.5f;
      const ZRay* zr_pair = A.zr + pair;            // record of sinogram s: [s * npair]
      const float* zt_pair = A.ztab ? A.ztab + pair : nullptr;  // entry e: [e * npair]
      for (int sg = 0; sg < A.nseg; sg++) {
        const GatherSeg S = segs[sg];
        int sl = (int)floorf((zmid_all - amid * S.dlo - S.z0) * S.dhi + 0.5f);
        sl = sl < 0 ? 0 : (sl > S.count - 1 ? S.count - 1 : sl);
        while (sl > 0) {
          const float2 z = s_zab[S.first + sl - 1];
          const float vz = z.y - z.x;
          if (fmaxf(z.x + vz * lo, z.x + vz * hi) < zlo_all) break;
          sl--;
        }
        for (; sl < S.count; sl++) {
          const int s = S.first + sl;
          const float2 zz = s_zab[s];
          const float vzz = zz.y - zz.x;
          const float za_lo = z_at(zz.x, vzz, lo), za_hi = z_at(zz.x, vzz, hi);
          const float zmax = fmaxf(za_lo, za_hi), zmin = fminf(za_lo, za_hi);
          if (zmax < zlo_all) continue;
          if (zmin > zhi_all) break;
          const ZRay z = load_zray(zr_pair + __umul24((unsigned)s, npair));
          // the table entry of each slice this sinogram can reach, loaded alongside the
          // record (its address does not depend on it) and used only when the ray
          // enters that slice by a z crossing of its own, which makes it an entry
          // zsetup wrote
          float ztv[ZV];
          bool hit[ZV];
#pragma unroll
          for (int q = 0; q < ZV; q++) {
            hit[q] = (q < nk) & !(zmax < zlo_t[q]) & !(zmin > zhi_t[q]);
            ztv[q] = 0.f;
          }
          if (zt_pair) {
            const int4 sl4 = s_slot[s];
#pragma unroll
            for (int q = 0; q < ZV; q++) {
              int e = kb + q + sl4.x;
              e = e < sl4.y ? sl4.y : (e > sl4.z ? sl4.z : e);
              if (hit[q]) ztv[q] = zt_pair[__umul24((unsigned)e, npair)];
            }
          }
          if (z.b == 0.f) continue;  // SiddonProjector.cpp:496
          const float az0 = z.az0, azu = z.azu;
          const int zk0r = z.k0, zkinc = z.kinc, zn = abs((int)z.k1 - z.k0);
#pragma unroll
          for (int q = 0; q < ZV; q++) {
            if (!hit[q]) continue;
            // the tests of gather_voxel, evaluated as one predicate (selects, not
            // branches); the term is added only when all hold
            const int k = kb + q;
            const int nc = __mul24(k - zk0r, zkinc);
            // (zkinc == 0 ? k == zk0r : nc >= 0), as lane-mask operations
            bool ok = (k == zk0r) | ((zkinc != 0) & (nc >= 0));
            const int m = na + nb + nc;
            ok &= m < tnxy + zn;
            if (!ok) continue;  // whole wavefronts often fail the count tests
            const bool hz = nc > 0;
            float Zt = ztv[q];
            if (ok & hz & !(zt_pair && nc <= z.zcount)) Zt = crossing<PREC>(az0, azu, nc - 1);
            const float Zp = hz ? Zt : -FLT_MAX;
            const float Zn = hz ? Zt + azu : az0;
            ok &= ((na <= 0) | (Xp <= Zn)) & ((nb <= 0) | (Yp <= Zn)) &
                  (!hz | ((Zp < Xn) & (Zp < Yn)));
            float ac = Xp > Yp ? Xp : Yp;
            ac = ac > Zp ? ac : Zp;
            ac = m > 0 ? ac : tamin;
            float an = Xn < Yn ? Xn : Yn;
            an = an < Zn ? an : Zn;
            const float w = (an - ac) * z.L;
            acc[q].add_if(ok, w * 1.0f * z.b);  // SiddonProjector.cpp:546
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < ZV; q++)
    if (q < nk) store_voxel(A, v0 + q * nxy, acc[q].value());
}

// Eight waves per SIMD for up to three voxels per lane: the kernel waits mostly on
// its record and table loads, and the two extra waves hide more of that latency than
// the few registers spilled at 64 VGPRs cost (3.29 -> 3.03 ms per C2 subset).
template <int PREC, int ZV, int GRID>
__global__ __launch_bounds__(kGBlock, ZV <= 3 ? 8 : 6) void gather_column_kernel(GatherArgs A) {
  extern __shared__ float2 s_zab[];
  const Geo& g = A.g;
  int4* s_slot = (int4*)((char*)s_zab + column_lds_off(g.nsino));
  GatherSeg* s_seg = (GatherSeg*)(s_slot + g.nsino);
  for (int q = threadIdx.x; q < g.nsino; q += kGBlock) {
    s_zab[q] = A.zab[q];
    if (A.zsoff) {
      const int so = A.zsoff[q], so1 = A.zsoff[q + 1], km = A.zsoff[g.nsino + 1 + q];
      s_slot[q] = make_int4(so - km, so, so1 - 1 < so ? so : so1 - 1, 0);
    }
  }
  for (int q = threadIdx.x; q < A.nseg; q += kGBlock) s_seg[q] = A.segs[q];
  __syncthreads();
  // a workgroup is a 16 x 16 tile of columns, ZV slices deep: the grid's own tile, or
  // the one the tile map gives this slot
  int tx = blockIdx.x, ty = blockIdx.y, tz = blockIdx.z;
  if (A.tile_map) {
    const int t = A.tile_map[blockIdx.x];
    if (t < 0) return;
    const int ntx = (g.nx + kTile - 1) / kTile, nty = (g.ny + kTile - 1) / kTile;
    tx = t % ntx;
    ty = (t / ntx) % nty;
    tz = t / (ntx * nty);
  }
  const int i = tx * kTile + (threadIdx.x % kTile);
  const int j = ty * kTile + (threadIdx.x / kTile);
  const int kb = tz * ZV;
  if (i >= g.nx || j >= g.ny || kb >= g.nz) return;
  gather_column<PREC, ZV, GRID>(A, s_zab, s_slot, s_seg, i, j, kb);
}

template <int DIMS, int PREC>
__global__ __launch_bounds__(kGBlock) void gather_kernel(GatherArgs A) {
  extern __shared__ float2 s_zab[];  // (P1.Z, P2.Z) of every sinogram
  const Geo& g = A.g;
  if (DIMS == 3) {
    int* s_soff = (int*)(s_zab + g.nsino);
    for (int q = threadIdx.x; q < g.nsino; q += kGBlock) s_zab[q] = A.zab[q];
    if (A.zsoff)
      for (int q = threadIdx.x; q <= 2 * g.nsino; q += kGBlock) s_soff[q] = A.zsoff[q];
    __syncthreads();
  }
  // Voxel of this lane.
  //   zlanes <= 1: a workgroup is a 16 x 16 tile of one slice (grid nx/16, ny/16, nz).
  //   zlanes = Z > 1: a wavefront holds 64 / Z voxel columns (neighbours along x) times
  //     Z consecutive slices; every lane of a column shares the column's candidate
  //     rays, so the transaxial tests agree across those lanes and only the axial
  //     part differs.  1D grid; workgroups are remapped so that each XCD takes one
  //     contiguous run of columns (round-robin dispatch puts block b on XCD b % 8).
  int i, j, k;
  if (A.zlanes <= 1) {
    i = blockIdx.x * kTile + (threadIdx.x % kTile);
    j = blockIdx.y * kTile + (threadIdx.x / kTile);
    k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
  } else {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned xcd = b % 8, per = nb / 8, rem = nb % 8;
    const unsigned lb = xcd * per + (xcd < rem ? xcd : rem) + b / 8;
    const int Z = A.zlanes, C = kWave / Z;
    const int nch = (g.nz + Z - 1) / Z;
    const long wave = ((long)lb * kGBlock + threadIdx.x) / kWave;
    const int lane = threadIdx.x % kWave;
    const int c = lane / Z, kk = lane % Z;
    const long q = (wave / nch) * C + c;
    k = (int)(wave % nch) * Z + kk;
    if (c >= C || k >= g.nz || q >= (long)g.nx * g.ny) return;
    i = (int)(q % g.nx);
    j = (int)(q / g.nx);
  }
  gather_voxel<DIMS, PREC>(A, s_zab, i, j, k);
}

// ---- grouped (multi-job) launches --------------------------------------------
#include "group_common.h"

__global__ __launch_bounds__(kGBlock) void zsetup_group_kernel(BatchHeader* __restrict__ hdr,
                                                               const ZsJob* __restrict__ jobs) {
  const int n = hdr->count[PH_ZSETUP];
  const long total = hdr->total[PH_ZSETUP];
  const int grab = grab_size(hdr);
  for (;;) {
    const long v0 = grab_vb(&hdr->next[PH_ZSETUP], grab);
    if (v0 >= total) break;
    for (long vb = v0; vb < v0 + grab && vb < total; vb++) {
      const ZsJob& J = jobs[find_job(jobs, n, vb)];
      const long idx = (vb - J.first) * kGBlock + threadIdx.x;
      if (idx < J.nrays)
        zsetup_one(J.g, J.rays, J.zab, J.sino, J.kind, J.nang_sub, J.subset, J.nsubsets, J.out,
                   J.nz_pair, J.ztab, J.zsoff, J.zblock, idx);
    }
  }
}

// Each virtual block is one 16 x 16 tile of one job (tile-major), of one slice in 2D
// and of kGroupZV slices in 3D (the column kernel's shape); 3D jobs reload the axial
// and crossing-slot tables into LDS when the workgroup moves to another job.
template <int DIMS>
__global__ __launch_bounds__(kGBlock) void gather_group_kernel(BatchHeader* __restrict__ hdr,
                                                               const GJob* __restrict__ jobs) {
  extern __shared__ float2 s_zab[];
  const int ph = DIMS == 2 ? PH_GATHER2 : PH_GATHER3;
  const int n = hdr->count[ph];
  const long total = hdr->total[ph];
  const int grab = grab_size(hdr);
  int loaded = -1;
  for (;;) {
    const long v0 = grab_vb(&hdr->next[ph], grab);
    if (v0 >= total) break;
    for (long vb = v0; vb < v0 + grab && vb < total; vb++) {
      const int jj = find_job(jobs, n, vb);
      const GJob& J = jobs[jj];
      const int ns = J.A.g.nsino;
      int4* s_slot = (int4*)((char*)s_zab + column_lds_off(ns));
      if (DIMS == 3 && loaded != jj) {
        __syncthreads();
        for (int q = threadIdx.x; q < ns; q += kGBlock) {
          s_zab[q] = J.A.zab[q];
          if (J.A.zsoff) {
            const int so = J.A.zsoff[q], so1 = J.A.zsoff[q + 1], km = J.A.zsoff[ns + 1 + q];
            s_slot[q] = make_int4(so - km, so, so1 - 1 < so ? so : so1 - 1, 0);
          }
        }
        __syncthreads();
        loaded = jj;
      }
      const long lb = vb - J.first;
      const int bx = (int)(lb % J.ntx);
      const int by = (int)((lb / J.ntx) % J.nty);
      const int kg = (int)(lb / ((long)J.ntx * J.nty));
      const int i = bx * kTile + (threadIdx.x % kTile);
      const int j = by * kTile + (threadIdx.x / kTile);
      if (i < J.A.g.nx && j < J.A.g.ny) {
        if constexpr (DIMS == 3)
          gather_column<APIRL_PREC_REFERENCE, kGroupZV>(J.A, s_zab, s_slot, J.A.segs, i, j,
                                                        kg * kGroupZV);
        else
          gather_voxel<DIMS, APIRL_PREC_REFERENCE>(J.A, s_zab, i, j, kg);
      }
    }
  }
}

inline unsigned ggrid(long n) { return (unsigned)((n + kGBlock - 1) / kGBlock); }

template <int PREC, int ZV>
void launch_gather_column(const GatherArgs& A, hipStream_t st) {
  const Geo& g = A.g;
  const dim3 grid = A.tile_map ? dim3((unsigned)A.tile_slots)
                               : dim3((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile,
                                      (g.nz + ZV - 1) / ZV);
  if (A.rlut)
    hipLaunchKernelGGL((gather_column_kernel<PREC, ZV, kGridLut>), grid, dim3(kGBlock),
                       column_lds_bytes(g.nsino, A.nseg), st, A);
  else
    hipLaunchKernelGGL((gather_column_kernel<PREC, ZV, kGridUniform>), grid, dim3(kGBlock),
                       column_lds_bytes(g.nsino, A.nseg), st, A);
}

template <int PREC>
void launch_gather3(const GatherArgs& A, int zlanes, int zvox, hipStream_t st) {
  const Geo& g = A.g;
  const size_t lds = gather_lds_bytes(g.nsino);
  if (PREC != APIRL_PREC_F64_ALPHA && PREC != APIRL_PREC_F64_SETUP && zlanes <= 1 &&
      zvox >= 1) {
    constexpr int P =
        (PREC == APIRL_PREC_F64_ALPHA || PREC == APIRL_PREC_F64_SETUP) ? APIRL_PREC_REFERENCE
                                                                       : PREC;
    if (zvox == 1) launch_gather_column<P, 1>(A, st);
    else if (zvox == 2) launch_gather_column<P, 2>(A, st);
    else if (zvox == 3) launch_gather_column<P, 3>(A, st);
    else launch_gather_column<P, 4>(A, st);
    return;
  }
  if (zlanes <= 1) {
    const dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, g.nz);
    hipLaunchKernelGGL((gather_kernel<3, PREC>), grid, dim3(kGBlock), lds, st, A);
    return;
  }
  const long ncol = (long)g.nx * g.ny;
  const int C = kWave / zlanes;
  const long nwaves = (ncol + C - 1) / C * ((g.nz + zlanes - 1) / zlanes);
  const long nblocks = (nwaves * kWave + kGBlock - 1) / kGBlock;
  hipLaunchKernelGGL((gather_kernel<3, PREC>), dim3((unsigned)nblocks), dim3(kGBlock), lds, st,
                     A);
}

template <int PREC>
void launch_gather2(const GatherArgs& A, hipStream_t st) {
  const Geo& g = A.g;
  const dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, g.nz);
  hipLaunchKernelGGL((gather_kernel<2, PREC>), grid, dim3(kGBlock), 0, st, A);
}


// ---- Tile-walk backprojector (3D, APIRL_PREC_REFERENCE) -----------------------------
//
// A wavefront owns an item: kTW x kTW voxel columns (a tile) times a slab of slices,
// summed in its own LDS.  For every ray (angle, radial bin) of the subset that crosses
// the tile, the wavefront finds once where the rays of all sinograms enter the tile
// (they share the transaxial path) and its lanes take the sinograms whose rays can
// reach the slab there, one ray per lane.  Each lane resumes the reference's Siddon
// loop (Siddon.cpp:310-340) at the tile's entry: the x / y crossings are the loop's
// accumulated float values from the plan's table, its z crossings come from the
// first one by the loop's own additions.  It walks to the tile's exit, adding
// (alpha - alpha_c) L b (SiddonProjector.cpp:546) to the voxels it visits inside the
// slab, while its step count stays below the list length |di| + |dj| + |dk|
// (:220, the dropped exit voxel), and only there: the voxels, weights and terms are
// the reference's floats, bit for bit.  Each voxel's terms are summed in double in
// the wavefront's order, which depends on the data only (one wavefront per item), so
// repeated runs are bit-identical, and rounded to float once.

constexpr int kTW = 8;      // item tile width in columns (kTW * kTW = one lane per column)
constexpr int kTWaves = 4;  // items (wavefronts) per workgroup

struct TileEntry {
  int i, j;      // column and row of the ray's first voxel in the tile
  int mxy;       // x and y crossings the loop has taken there
  float ce;      // alpha_c there: the crossing just taken, or alpha_min (first voxel)
  float Xn, Yn;  // the next x and y crossings (the loop's accumulated values)
  float aout;    // alpha near the tile's exit (sinogram candidates only)
};

// A candidate ray of a tile, found once per subset by tile_rays_kernel for every
// slab of the tile: its entry and the length of its transaxial path there.
struct alignas(16) TileRay {
  TileEntry E;
  int pair;      // ka * nr + r, or -1: no ray (or one that misses the tile)
  int ncol;      // columns on its path through the tile
  int nxy;       // the ray's transaxial list length |di| + |dj|
  int pad[2];
};
static_assert(sizeof(TileRay) == 48, "TileRay is three 16-byte loads");

// The transaxial part of a candidate ray's walk through its tile, the same for every
// sinogram: the columns it visits in order and, for each, the alpha at which the
// loop leaves it by an x or y crossing (the smaller accumulated crossing there; x
// before y on ties, Siddon.cpp:318-331).  At most 2 kTW - 1 columns.
constexpr int kTPath = 2 * kTW;
struct alignas(16) TilePath {
  float ae[kTPath];
  unsigned char col[kTPath];  // (j - tj0) kTW + (i - ti0)
};
static_assert(sizeof(TilePath) == 80, "TilePath is five 16-byte loads");

// Crossings p after which index i0 + inc p enters [lo, hi] (pin) and first lies
// beyond it (pout); false when it never lies in it.
__device__ __forceinline__ bool axis_span(int i0, int inc, int lo, int hi, int& pin, int& pout) {
  if (inc == 0) {
    pin = 0;
    pout = INT_MAX;
    return i0 >= lo && i0 <= hi;
  }
  if (inc > 0) {
    pin = lo - i0 > 0 ? lo - i0 : 0;
    pout = hi - i0 + 1;
    return i0 <= hi;
  }
  pin = i0 - hi > 0 ? i0 - hi : 0;
  pout = i0 - lo + 1;
  return i0 >= lo;
}

// Crossing p >= 1 along one axis: the table's value for index c = i0 + inc p, or, for
// the one index past the image that an item at its edge asks for, the loop's next
// addition to the previous crossing.
__device__ __forceinline__ float axis_cross(const float* tab, int n, int i0, int inc, float a0,
                                            float au, int p) {
  const int c = i0 + inc * p;
  if (c >= 0 && c < n) return tab[c];
  return p == 1 ? a0 : tab[c - inc] + au;
}

// Number of crossings p >= 1 whose value is <= v (strict: < v), given that crossing
// plo (when plo >= 1) is and that the answer is taken as phi once it reaches phi:
// the linear estimate, corrected against the exact values.
__device__ __forceinline__ int axis_count(const float* tab, int n, int i0, int inc, float a0,
                                          float au, int plo, int phi, float v, bool strict) {
  const float e = (v - a0) / au + 1.f;
  int p = !(e > (float)plo) ? plo : (e < (float)phi ? (int)e : phi);
  while (p > plo) {
    const float x = axis_cross(tab, n, i0, inc, a0, au, p);
    if (strict ? x < v : x <= v) break;
    p--;
  }
  while (p < phi) {
    const float x = axis_cross(tab, n, i0, inc, a0, au, p + 1);
    if (!(strict ? x < v : x <= v)) break;
    p++;
  }
  return p;
}

// Where ray t enters the tile [ti0, ti1] x [tj0, tj1]: the loop's state after the
// later of the crossings that bring i and j into range (an x crossing at the same
// alpha as a y crossing is taken first, Siddon.cpp:318-331); false when the ray's
// path misses the tile.
__device__ __forceinline__ bool tile_entry(const Geo& g, const RayXY& t, const float* xy, int ti0,
                                           int ti1, int tj0, int tj1, TileEntry& E) {
  const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
  int pin, pout, qin, qout;
  if (!axis_span(t.i0, iinc, ti0, ti1, pin, pout) || !axis_span(t.j0, jinc, tj0, tj1, qin, qout))
    return false;
  const float* yt = xy + g.nx;
  int p = 0, q = 0;
  float ce = t.amin;
  if (pin > 0 || qin > 0) {
    const float Xin = pin > 0 ? axis_cross(xy, g.nx, t.i0, iinc, t.ax, t.axu, pin) : -FLT_MAX;
    const float Yin = qin > 0 ? axis_cross(yt, g.ny, t.j0, jinc, t.ay, t.ayu, qin) : -FLT_MAX;
    if (Xin <= Yin) {  // entered by y crossing qin; x crossings up to its alpha are taken
      q = qin;
      ce = Yin;
      if (iinc != 0) {
        p = axis_count(xy, g.nx, t.i0, iinc, t.ax, t.axu, pin, pout, Yin, false);
        if (p >= pout) return false;
      }
    } else {  // entered by x crossing pin; y crossings below its alpha are taken
      p = pin;
      ce = Xin;
      if (jinc != 0) {
        q = axis_count(yt, g.ny, t.j0, jinc, t.ay, t.ayu, qin, qout, Xin, true);
        if (q >= qout) return false;
      }
    }
  }
  E.i = t.i0 + iinc * p;
  E.j = t.j0 + jinc * q;
  E.mxy = p + q;
  E.ce = ce;
  E.Xn = (iinc == 0 || p == 0) ? t.ax : xy[E.i] + t.axu;
  E.Yn = (jinc == 0 || q == 0) ? t.ay : yt[E.j] + t.ayu;
  const float xo = iinc == 0 ? FLT_MAX : t.ax + (float)(pout - 1) * t.axu;
  const float yo = jinc == 0 ? FLT_MAX : t.ay + (float)(qout - 1) * t.ayu;
  E.aout = fminf(t.amax, fminf(xo, yo));
  return true;
}

// Radial bins of angle a whose LORs can cross a tile centred at (xc, yc) with
// half-widths (hx, hy): candidate_bins for a rectangle.
__device__ __forceinline__ void tile_bins(const GatherArgs& A, int a, float xc, float yc, float hx,
                                          float hy, float inv_dr, int& lo0, int& hi0, int& lo1,
                                          int& hi1) {
  const Geo& g = A.g;
  const float2 cs = A.angcs[a];
  const float sc = xc * cs.x + yc * cs.y;
  const float hw = hx * fabsf(cs.x) + hy * fabsf(cs.y) + g.rmargin;
  if (!grid_lut<kGridAny>(A)) {
    int jlo = (int)ceilf((sc - hw - g.r0) * inv_dr);
    int jhi = (int)floorf((sc + hw - g.r0) * inv_dr);
    lo0 = lo1 = jlo < 0 ? 0 : jlo;
    hi0 = hi1 = jhi > g.nr - 1 ? g.nr - 1 : jhi;
    return;
  }
  lut_bins(A, sc, hw, lo0, hi0);
  const float2 co = A.angcs_odd[a];
  const float so = xc * co.x + yc * co.y;
  const float hwo = hx * fabsf(co.x) + hy * fabsf(co.y) + g.rmargin;
  lut_bins(A, so, hwo, lo1, hi1);
}

// One lane: the ray of record z from the tile's entry E to its exit, its terms added
// to the item's sums acc[column * zs + slice - kz0] for slices kz0..kz1.  One loop
// iteration is one step of the reference's loop, its branch taken by selects.
__device__ __forceinline__ void tile_lane(const TileRay& R, const float* pae,
                                          const unsigned char* pcol, const ZRay& z, int kz0,
                                          int kz1, int zs, double* acc) {
  const TileEntry& E = R.E;
  const int nxy = R.nxy, ncol = R.ncol;
  const int kinc = z.kinc;
  const float azu = z.azu, L = z.L, b = z.b;
  const int N = nxy + abs((int)z.k1 - z.k0);
  int m = E.mxy, k = z.k0;
  float Zn = z.az0;
  // z crossings taken before the entry: those below its alpha (a z crossing at the
  // alpha of an x or y crossing is taken after it, Siddon.cpp:318-339), by the
  // loop's own additions, four at a time
  if (m > 0 && kinc != 0) {
    while (m < N) {
      const float z1 = Zn + azu, z2 = z1 + azu, z3 = z2 + azu;
      if (z3 < E.ce) {
        Zn = z3 + azu;
        k += 4 * kinc;
        m += 4;
        continue;
      }
      const int c = (Zn < E.ce ? 1 : 0) + (z1 < E.ce ? 1 : 0) + (z2 < E.ce ? 1 : 0);
      Zn = c == 0 ? Zn : (c == 1 ? z1 : (c == 2 ? z2 : z3));
      k += c * kinc;
      m += c;
      break;
    }
  }
  // slice relative to the slab and its end in the direction of travel (a lane that
  // reaches it adds nothing more); the column comes from the path, whose x / y steps
  // are the same for every lane
  const int nzs = kz1 - kz0 + 1;
  int kl = k - kz0;
  const int kend = kinc > 0 ? nzs : -1;
  if (kinc == 0 ? (unsigned)kl >= (unsigned)nzs : (kinc > 0 ? kl >= nzs : kl < 0)) return;
  int q = 0;
  float ae = pae[0];
  int cidx = pcol[0] * zs;
  float ac = E.ce;
  for (int left = N - m; left > 0; left--) {
    // the next column's entry, read before this step's add so that the wait for it
    // does not include the add (entry q + 1 < kTPath always exists)
    const float ae1 = pae[q + 1];
    const int c1 = pcol[q + 1];
    const bool zst = Zn < ae;  // Siddon.cpp:318-339: x, then y, then z on ties
    const float an = zst ? Zn : ae;
    if ((unsigned)kl < (unsigned)nzs) {
      const float term = (an - ac) * L * b;  // the reference's float term (geomFactor 1)
      __hip_atomic_fetch_add(acc + cidx + kl, (double)term, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    ac = an;
    Zn = zst ? Zn + azu : Zn;
    kl += zst ? kinc : 0;
    q += zst ? 0 : 1;
    if ((q == ncol) | (kl == kend)) break;
    ae = zst ? ae : ae1;
    cidx = zst ? cidx : c1 * zs;
  }
}

// Orders a wavefront's LDS accesses before and after this point (lanes read what
// other lanes of the wavefront wrote).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One in-flight ray of a wavefront, in LDS: its entry and path (the lanes walking
// its sinograms read both).
struct alignas(16) TileSlot {
  TileRay R;
  TilePath P;
};
static_assert(sizeof(TileSlot) == 128, "TileSlot is 32 words");
constexpr int kTSlots = 8;    // rays in flight per wavefront
constexpr int kTRing = 128;   // pending (ray, sinogram) entries per wavefront

// Bytes of one wavefront's LDS: the item's sums (double: LDS adds of doubles run
// about 20 times faster than of floats here, and sum more exactly), the rays in
// flight, and the ring of pending (ray, sinogram) entries.
__host__ __device__ inline int tile_wave_bytes(int zs, int nsino) {
  (void)nsino;
  return kTW * kTW * zs * 8 + kTSlots * (int)sizeof(TileSlot) + kTRing * 4;
}


// Candidate ray c of angle ka for tile `tile` (slot (tile * nang_sub + ka) * maxc + c):
// the c-th bin of the angle's candidate range, kept when its bins are not all zero,
// its LOR is valid and its path enters the tile.  The slots of one (tile, angle) are
// in bin order.
__global__ __launch_bounds__(kGBlock) void tile_rays_kernel(GatherArgs A, int maxc, long nslots,
                                                            TileRay* __restrict__ out,
                                                            TilePath* __restrict__ paths) {
  const long idx = (long)blockIdx.x * kGBlock + threadIdx.x;
  if (idx >= nslots) return;
  const Geo& g = A.g;
  const int c = (int)(idx % maxc);
  const long t2 = idx / maxc;
  const int ka = (int)(t2 % A.nang_sub);
  const int tile = (int)(t2 / A.nang_sub);
  const int ntx = (g.nx + kTW - 1) / kTW;
  const int ti0 = (tile % ntx) * kTW, tj0 = (tile / ntx) * kTW;
  const int ti1 = (ti0 + kTW < g.nx ? ti0 + kTW : g.nx) - 1;
  const int tj1 = (tj0 + kTW < g.ny ? tj0 + kTW : g.ny) - 1;
  const float xt = -g.R + g.dx * 0.5f * (float)(ti0 + ti1 + 1);
  const float yt = -g.R + g.dy * 0.5f * (float)(tj0 + tj1 + 1);
  const float hx = 0.5f * g.dx * (float)(ti1 - ti0 + 1);
  const float hy = 0.5f * g.dy * (float)(tj1 - tj0 + 1);
  const int a = A.subset + A.nsubsets * ka;
  int lo0, hi0, lo1, hi1;
  tile_bins(A, a, xt, yt, hx, hy, 1.0f / g.dr, lo0, hi0, lo1, hi1);
  const int jlo = lo0 < lo1 ? lo0 : lo1, jhi = hi0 > hi1 ? hi0 : hi1;
  const int rr = jlo + c;
  TileRay R;
  R.pair = -1;
  if (rr <= jhi && (!grid_lut<kGridAny>(A) || in_parity_range(rr, lo0, hi0, lo1, hi1))) {
    const int pair = ka * g.nr + rr;
    const RayXY t = A.rays[(long)a * g.nr + rr];
    if ((t.flags & RAY_VALID) &&
        tile_entry(g, t, A.xytab + ((long)a * g.nr + rr) * (g.nx + g.ny), ti0, ti1, tj0, tj1,
                   R.E)) {
      R.pair = pair;
      R.nxy = t.nxy;
      // the path: the loop's x / y steps from the entry until it leaves the tile
      TilePath P;
      const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
      float Xn = R.E.Xn, Yn = R.E.Yn;
      int i = R.E.i, j = R.E.j, q = 0;
      while (q < kTPath) {
        P.col[q] = (unsigned char)((j - tj0) * kTW + (i - ti0));
        P.ae[q] = Xn < Yn ? Xn : Yn;
        q++;
        if (Xn <= Yn) {
          i += iinc;
          Xn = Xn + t.axu;
        } else {
          j += jinc;
          Yn = Yn + t.ayu;
        }
        if (i < ti0 || i > ti1 || j < tj0 || j > tj1) break;
      }
      for (int r = q; r < kTPath; r++) {
        P.col[r] = 0;
        P.ae[r] = 0.f;
      }
      R.ncol = q;
      paths[idx] = P;
    }
  }
  out[idx] = R;
}

// The candidate rays of each tile that enter it, moved to the front of its slots in
// slot order (one wavefront per tile); count[tile] says how many.
__global__ __launch_bounds__(kGBlock) void tile_compact_kernel(TileRay* __restrict__ trays,
                                                               TilePath* __restrict__ paths,
                                                               int nslot, int ntiles,
                                                               int* __restrict__ count) {
  const int tile = blockIdx.x * kTWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  if (tile >= ntiles) return;
  TileRay* tr = trays + (long)tile * nslot;
  TilePath* tp = paths + (long)tile * nslot;
  int n = 0;
  for (int c0 = 0; c0 < nslot; c0 += kWave) {
    TileRay R;
    TilePath P;
    R.pair = -1;
    if (c0 + lane < nslot) R = tr[c0 + lane];
    const bool keep = R.pair >= 0;
    if (keep) P = tp[c0 + lane];
    const unsigned long long bal = __ballot(keep);
    __builtin_amdgcn_wave_barrier();  // every lane has read its slot before any is written
    if (keep) {
      const int d = n + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
      tr[d] = R;
      tp[d] = P;
    }
    n += __popcll(bal);
  }
  if (lane == 0) count[tile] = n;
}

// The tile backprojector's axial records, out[(ka nr + r) nsino + s] (zray_setup and
// the bin value), through a 32 (r) x 32 (s) LDS tile of the bin values so that the
// sinogram is read along r and the records are written along s; nz_pair[ka nr + r]
// set where a valid LOR's bin is non-zero (one atomic per 32 sinograms).
__global__ __launch_bounds__(kGBlock) void tile_zsetup_kernel(GatherArgs A,
                                                              ZRay* __restrict__ out,
                                                              int* __restrict__ nz_pair) {
  __shared__ float tb[32][33];
  const Geo& g = A.g;
  const int r0 = blockIdx.x * 32, ka = blockIdx.y, s0 = blockIdx.z * 32;
  const int a = A.subset + A.nsubsets * ka;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int q = ty; q < 32; q += 8) {
    const int s = s0 + q, rr = r0 + tx;
    float b = 1.f;
    if (s < g.nsino && rr < g.nr) {
      if (A.kind == BP_PLAIN) b = A.sino[((long)s * A.nang_sub + ka) * g.nr + rr];
      else if (A.kind == BP_MULT) b = A.sino[((long)s * g.nang + a) * g.nr + rr];
    }
    tb[q][tx] = b;
  }
  __syncthreads();
  const int s = s0 + tx;
  const float2 zz = s < g.nsino ? A.zab[s] : make_float2(0.f, 0.f);
  for (int q = ty; q < 32; q += 8) {
    const int rr = r0 + q;  // uniform over each half wavefront
    bool nz = false;
    if (rr < g.nr && s < g.nsino) {
      const RayXY R = A.rays[(long)a * g.nr + rr];
      ZRay z;
      zray_setup(g, R, zz.x, zz.y, z);
      z.b = tb[tx][q];
      z.zcount = 0;
      z.pad = 0;
      out[((long)ka * g.nr + rr) * g.nsino + s] = z;
      nz = z.b != 0.f && (R.flags & RAY_VALID);
    }
    const unsigned long long bal = __ballot(nz);
    const unsigned half = (unsigned)(bal >> (threadIdx.x & 32));
    if (tx == 0 && half != 0u) atomicOr(&nz_pair[(long)ka * g.nr + rr], 1);
  }
}

// Per tile, the positions (in trays order) of its rays with a non-zero bin, live[tile *
// nslot + c] for c < nlive[tile] (one wavefront per tile).
__global__ __launch_bounds__(kGBlock) void tile_live_kernel(const TileRay* __restrict__ trays,
                                                            const int* __restrict__ count,
                                                            const int* __restrict__ nz_pair,
                                                            int nslot, int ntiles,
                                                            int* __restrict__ live,
                                                            int* __restrict__ nlive) {
  const int tile = blockIdx.x * kTWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  if (tile >= ntiles) return;
  const TileRay* tr = trays + (long)tile * nslot;
  int* lv = live + (long)tile * nslot;
  const int nc = count[tile];
  int n = 0;
  for (int c0 = 0; c0 < nc; c0 += kWave) {
    const int c = c0 + lane;
    const bool keep = c < nc && nz_pair[tr[c].pair] != 0;
    const unsigned long long bal = __ballot(keep);
    if (keep)
      lv[n + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                       __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0))] = c;
    n += __popcll(bal);
  }
  if (lane == 0) nlive[tile] = n;
}

// The sinograms whose ray (tile entry E) can reach the slab [zlo, zhi] (widened)
// inside the tile, in segment-run order: per run, the indices whose mean z lies
// within the slab widened by the run's slopes over the tile's alpha range (z = zmean
// + (zb - za)(alpha - 1/2)) and its deviation from the line (one run per lane); of
// those, the sinograms whose ray's z over that alpha range meets the slab.  The
// wavefront calls emit(keep, s) together for each group of up to 64 tested indices.
template <class Emit>
__device__ __forceinline__ void tile_list_ray(const GatherArgs& A, const TileEntry& E, float zlo,
                                              float zhi, const float2* s_zab, int lane,
                                              Emit&& emit) {
  const float ulo = E.ce - 0.5f, uhi = E.aout - 0.5f;
  for (int sb = 0; sb < A.nseg; sb += kWave) {
    int s0 = 0, cnt = 0;
    if (sb + lane < A.nseg) {
      const GatherSeg S = A.segs[sb + lane];
      const float c1 = S.dmin * ulo, c2 = S.dmin * uhi, c3 = S.dmax * ulo, c4 = S.dmax * uhi;
      const float dlo = fminf(fminf(c1, c2), fminf(c3, c4));
      const float dhi = fmaxf(fmaxf(c1, c2), fmaxf(c3, c4));
      const float lf = (zlo - dhi - S.devz - 0.01f - S.z0) * S.dhi - 1.f;
      const float hf = (zhi - dlo + S.devz + 0.01f - S.z0) * S.dhi + 1.f;
      const int l0 = lf > 0.f ? (int)lf : 0;
      const int l1 = hf < (float)(S.count - 1) ? (int)hf : S.count - 1;
      s0 = S.first + l0;
      cnt = l1 >= l0 ? l1 - l0 + 1 : 0;
    }
    const int ns = A.nseg - sb < kWave ? A.nseg - sb : kWave;
    for (int sg = 0; sg < ns; sg++) {
      const int cs = __builtin_amdgcn_readlane(cnt, sg);
      if (cs == 0) continue;
      const int f = __builtin_amdgcn_readlane(s0, sg);
      for (int l0 = 0; l0 < cs; l0 += kWave) {
        const int s = f + l0 + lane;
        bool keep = false;
        if (l0 + lane < cs) {
          const float2 zz = s_zab[s];
          const float vz = zz.y - zz.x;
          const float z1 = z_at(zz.x, vz, E.ce), z2 = z_at(zz.x, vz, E.aout);
          keep = !(fmaxf(z1, z2) < zlo) && !(fminf(z1, z2) > zhi);
        }
        emit(keep, s);
      }
    }
  }
}

// The slab's z range widened by the listing margin.
__device__ __forceinline__ void tile_slab_z(const Geo& g, int kz0, int kz1, float& zlo,
                                            float& zhi) {
  zlo = g.offz + kz0 * g.dz - kZMargin;
  zhi = g.offz + (kz1 + 1) * g.dz + kZMargin;
}

// The listed sinograms of every (item, candidate ray) kept for a subset, one wavefront
// per item, rays in the tile's order.  COUNT: loff[item * (nslot + 1) + c] = entries
// of the item before ray c (c up to the tile's ray count) and tot[item] = its total.
// Otherwise the sinograms (16 bits each) go to list[base[item] + ...].
template <bool COUNT>
__global__ __launch_bounds__(kGBlock) void tile_list_kernel(
    GatherArgs A, int zs, int nslab, int nitems, const TileRay* __restrict__ trays,
    const int* __restrict__ tcount, int nslot, int* __restrict__ loff, int* __restrict__ tot,
    const long long* __restrict__ base, unsigned short* __restrict__ list) {
  extern __shared__ float2 lzab[];
  const Geo& g = A.g;
  for (int q = threadIdx.x; q < g.nsino; q += kGBlock) lzab[q] = A.zab[q];
  __syncthreads();
  const int item = blockIdx.x * kTWaves + (threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  if (item >= nitems) return;
  const int slab = item % nslab, tile = item / nslab;
  const int kz0 = slab * zs, kz1 = (kz0 + zs < g.nz ? kz0 + zs : g.nz) - 1;
  float zlo, zhi;
  tile_slab_z(g, kz0, kz1, zlo, zhi);
  const TileRay* tr = trays + (long)tile * nslot;
  const int nray = tcount[tile];
  int n = 0;
  unsigned short* out = COUNT ? nullptr : list + base[item];
  for (int c = 0; c < nray; c++) {
    if (COUNT && lane == 0) loff[(long)item * (nslot + 1) + c] = n;
    const TileEntry E = tr[c].E;
    tile_list_ray(A, E, zlo, zhi, lzab, lane, [&](bool keep, int s) {
      const unsigned long long bal = __ballot(keep);
      if (!COUNT && keep)
        out[n + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0))] =
            (unsigned short)s;
      n += __popcll(bal);
    });
  }
  if (COUNT && lane == 0) {
    loff[(long)item * (nslot + 1) + nray] = n;
    tot[item] = n;
  }
}

// base[i] = sum of tot[0 .. i), for i = 0 .. n, in one workgroup.
__global__ __launch_bounds__(kGBlock) void tile_scan_kernel(const int* __restrict__ tot, int n,
                                                            long long* __restrict__ base) {
  __shared__ long long part[kGBlock];
  const int per = (n + kGBlock - 1) / kGBlock;
  const int b = threadIdx.x * per, e = b + per < n ? b + per : n;
  long long sum = 0;
  for (int i = b; i < e; i++) sum += tot[i];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long run = 0;
    for (int t = 0; t < kGBlock; t++) {
      const long long v = part[t];
      part[t] = run;
      run += v;
    }
    base[n] = run;
  }
  __syncthreads();
  long long run = part[threadIdx.x];
  for (int i = b; i < e; i++) {
    base[i] = run;
    run += tot[i];
  }
}

// Items (tile, slab) are taken from a queue in the order of order[] (the tiles
// nearest the image centre first, whose rays are the most), counter at *next.  The
// rays entering tile t are trays[t * nslot + c]; those with a non-zero bin are at the
// positions live[t * nslot + c'], c' < nlive[t].  The (ray, sinogram) pairs to walk
// go through a ring in LDS: a pass takes 64 of them as soon as there are, so a pass
// carries the sinograms of several rays.
// LISTED: each (item, ray)'s sinograms are read from the lists tile_list_kernel keeps
// (loff, base, list) instead of being listed here.
template <bool LISTED>
__global__ __launch_bounds__(kGBlock) void tile_bp_kernel(GatherArgs A, int zs, int nslab,
                                                          const int* __restrict__ order,
                                                          int nitems, int* __restrict__ next,
                                                          const TileRay* __restrict__ trays,
                                                          const TilePath* __restrict__ tpaths,
                                                          int nslot,
                                                          const int* __restrict__ live,
                                                          const int* __restrict__ nlive,
                                                          const int* __restrict__ loff,
                                                          const long long* __restrict__ lbase,
                                                          const unsigned short* __restrict__ list) {
  extern __shared__ double tlds[];
  const Geo& g = A.g;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & (kWave - 1);
  const int ntx = (g.nx + kTW - 1) / kTW;
  double* acc = (double*)((char*)tlds + (size_t)w * tile_wave_bytes(zs, g.nsino));
  TileSlot* slots = (TileSlot*)(acc + kTW * kTW * zs);
  unsigned* ring = (unsigned*)(slots + kTSlots);
  // the axial table, shared by the workgroup's wavefronts, for the listing (the kept
  // lists need none: their launch leaves it out of the LDS)
  float2* s_zab = (float2*)((char*)tlds + (size_t)kTWaves * tile_wave_bytes(zs, g.nsino));
  if (!LISTED)
    for (int q = threadIdx.x; q < g.nsino; q += kGBlock) s_zab[q] = A.zab[q];
  __syncthreads();
  const long nxy = (long)g.nx * g.ny;
  const int il = lane % kTW, jl = lane / kTW;
  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(next, 1);
    q = __builtin_amdgcn_readfirstlane(q);
    if (q >= nitems) return;
    const int item = order[q];
    const int slab = item % nslab, tile = item / nslab;
    const int ti0 = (tile % ntx) * kTW, tj0 = (tile / ntx) * kTW;
    const int ti1 = (ti0 + kTW < g.nx ? ti0 + kTW : g.nx) - 1;
    const int tj1 = (tj0 + kTW < g.ny ? tj0 + kTW : g.ny) - 1;
    const int kz0 = slab * zs, kz1 = (kz0 + zs < g.nz ? kz0 + zs : g.nz) - 1;
    const bool incol = ti0 + il <= ti1 && tj0 + jl <= tj1;
    const long v0 = (long)(tj0 + jl) * g.nx + ti0 + il;
    wave_lds_sync();  // the previous item's sums are read out
    for (int kl = 0; kl <= kz1 - kz0; kl++)
      acc[lane * zs + kl] = (incol && !A.upd_sens) ? (double)A.img[(kz0 + kl) * nxy + v0] : 0.0;
    wave_lds_sync();
    float zlo, zhi;
    tile_slab_z(g, kz0, kz1, zlo, zhi);
    const long long ibase = LISTED ? lbase[item] : 0;
    const TileRay* tr = trays + (long)tile * nslot;
    const unsigned* trw = (const unsigned*)tr;
    const unsigned* tpw = (const unsigned*)(tpaths + (long)tile * nslot);
    const int* lv = live + (long)tile * nslot;
    const int nray = nlive[tile];
    int head = 0, pend = 0;  // ring entries (ray << 16 | sinogram), head .. head + pend - 1
    int lvc = 0;             // live positions c0 .. c0 + 63, one per lane
    // one pass: lanes l < k walk entry head + l
    auto run_pass = [&](int k) {
      wave_lds_sync();
      if (lane < k) {
        const unsigned e = ring[(head + lane) & (kTRing - 1)];
        const TileSlot& S = slots[(e >> 16) & (kTSlots - 1)];
        const ZRay z = load_zray(A.zr + (long)S.R.pair * g.nsino + (e & 0xffffu));
        if (z.b != 0.f)  // SiddonProjector.cpp:496
          tile_lane(S.R, S.P.ae, S.P.col, z, kz0, kz1, zs, acc);
      }
      head = (head + k) & (kTRing - 1);
      pend -= k;
    };
    for (int c = 0; c < nray; c++) {
      // the slot of ray c is free once no pending entry belongs to ray c - kTSlots
      while (pend > 0) {
        wave_lds_sync();
        if ((int)(ring[head] >> 16) + kTSlots > c) break;
        run_pass(pend < kWave ? pend : kWave);
      }
      wave_lds_sync();
      if ((c & (kWave - 1)) == 0) lvc = c + lane < nray ? lv[c + lane] : 0;
      const int rc = __builtin_amdgcn_readlane(lvc, c & (kWave - 1));  // in the tile's list
      if (lane < (int)(sizeof(TileRay) / 4))
        ((unsigned*)&slots[c & (kTSlots - 1)].R)[lane] =
            trw[(long)rc * (sizeof(TileRay) / 4) + lane];
      else if (lane < (int)(sizeof(TileSlot) / 4))
        ((unsigned*)&slots[c & (kTSlots - 1)].R)[lane] =
            tpw[(long)rc * (sizeof(TilePath) / 4) + lane - (int)(sizeof(TileRay) / 4)];
      // the sinograms of ray c that reach the slab inside the tile go to the ring
      auto emit = [&](bool keep, int s) {
        const unsigned long long bal = __ballot(keep);
        if (keep)
          ring[(head + pend +
                __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0))) &
               (kTRing - 1)] = ((unsigned)c << 16) | (unsigned)s;
        pend += __popcll(bal);
        if (pend >= kWave) run_pass(kWave);
      };
      if (LISTED) {
        const int* lo = loff + (long)item * (nslot + 1) + rc;
        const long long b0 = ibase + lo[0], b1 = ibase + lo[1];
        for (long long l0 = b0; l0 < b1; l0 += kWave) {
          const bool keep = l0 + lane < b1;
          emit(keep, keep ? (int)list[l0 + lane] : 0);
        }
      } else {
        tile_list_ray(A, tr[rc].E, zlo, zhi, s_zab, lane, emit);
      }
    }
    while (pend > 0) run_pass(pend < kWave ? pend : kWave);
    wave_lds_sync();
    if (incol)
      for (int kl = 0; kl <= kz1 - kz0; kl++)
        store_voxel(A, (kz0 + kl) * nxy + v0, (float)acc[lane * zs + kl]);
  }
}

}  // namespace

hipError_t launch_backproject_gather(const apirl_plan* p, const float* sino, float* img,
                                     int subset, int nsubsets, int kind, int precision,
                                     void* scratch, size_t scratch_bytes, hipStream_t st,
                                     const EmFuse* fuse) {
  if (kind != BP_PLAIN && kind != BP_ONES && kind != BP_MULT) return hipErrorNotSupported;
  if (precision < APIRL_PREC_REFERENCE || precision > APIRL_PREC_F64_SETUP)
    return hipErrorInvalidValue;
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  const long nvox = (long)g.nx * g.ny * g.nz;
  if (nrays == 0 || nvox == 0) return hipSuccess;
  if (g.dims == 3 && scratch_bytes < gather_scratch_bytes(p, subset, nsubsets))
    return hipErrorInvalidValue;
  GatherArgs A;
  ZRay* zr = nullptr;
  int* nzp = nullptr;
  gather_args_for(p, subset, nsubsets, kind, sino, img, g.dims == 3 ? scratch : nullptr, A, zr,
                  nzp);
  A.zlanes = g.dims == 3 ? p->gather_zlanes : 1;
  if (fuse) {
    A.upd_sens = fuse->sens;
    A.upd_thr = fuse->thr;
    A.upd_strict = fuse->strict;
  }
  if (g.dims == 3) {
    // APIRL_PREC_F64_SETUP with its tables (the rays' double transaxial setup, their
    // entry crossings, and zsetup's double axial setup): the column kernel, whose
    // terms are then those of the one-voxel F64_SETUP gather, summed as KAHAN
    const LorForm& F = p->forms[p->scanner];
    const bool f64s = precision == APIRL_PREC_F64_SETUP && A.zlanes <= 1 && p->gather_zvox >= 1 &&
                      F.rays_f64s && F.xytab_f64s;
    if (f64s) {
      A.rays = F.rays_f64s;
      A.xytab = F.xytab_f64s;
    }
    const long npair = (long)nang_sub * g.nr;
    hipLaunchKernelGGL(clear_int_kernel, dim3(ggrid(npair)), dim3(kGBlock), 0, st, nzp, npair);
    hipLaunchKernelGGL(zsetup_kernel, dim3(ggrid(nrays)), dim3(kGBlock), 0, st, g, p->d_rays,
                       p->d_zab, sino, kind, nang_sub, subset, nsubsets, nrays, zr, nzp, A.ztab,
                       A.zsoff, A.zblock, f64s ? 1 : 0);
    const int zv = p->gather_zvox;
    if (precision == APIRL_PREC_REFERENCE)
      launch_gather3<APIRL_PREC_REFERENCE>(A, A.zlanes, zv, st);
    else if (precision == APIRL_PREC_KAHAN || f64s)
      launch_gather3<APIRL_PREC_KAHAN>(A, A.zlanes, zv, st);
    else if (precision == APIRL_PREC_F64_SETUP)
      launch_gather3<APIRL_PREC_F64_SETUP>(A, A.zlanes, zv, st);
    else launch_gather3<APIRL_PREC_F64_ALPHA>(A, A.zlanes, zv, st);
  } else {
    if (precision == APIRL_PREC_REFERENCE) launch_gather2<APIRL_PREC_REFERENCE>(A, st);
    else if (precision == APIRL_PREC_KAHAN || precision == APIRL_PREC_F64_SETUP)
      launch_gather2<APIRL_PREC_KAHAN>(A, st);
    else launch_gather2<APIRL_PREC_F64_ALPHA>(A, st);
  }
  return hipGetLastError();
}

// Slab depth of the tile backprojector's items: the plan's tuning, or about 5
// slices (shallow items keep few lanes on one voxel per LDS add and leave room for
// more wavefronts per SIMD; with the kept lists, 1.26 ms per 12-angle subset of the
// 256x256x81 / 344x252x837 workload at 5 slices, 1.31 at 4 and 1.32 at 6).
int tile_slab_depth(const apirl_plan* p) {
  const int nz = p->geo.nz;
  if (p->tile_zs > 0) return p->tile_zs < nz ? p->tile_zs : nz;
  const int nslab = (nz + 4) / 5;
  return (nz + nslab - 1) / nslab;
}

size_t tile_lds_bytes(const apirl_plan* p, bool listed) {
  if (listed) return (size_t)kTWaves * tile_wave_bytes(tile_slab_depth(p), p->geo.nsino);
  return (size_t)kTWaves * tile_wave_bytes(tile_slab_depth(p), p->geo.nsino) +
         (size_t)p->geo.nsino * sizeof(float2);
}

// Upper bound on the candidate bins of one (tile, angle): the bins within the tile's
// projected half-width (plus the search margin, and one lookup-table step on each
// side on a non-uniform grid) of its centre, for each parity.
int tile_max_candidates(const apirl_plan* p) {
  const Geo& g = p->geo;
  double drmin = g.dr > 0 ? g.dr : 1.0;
  for (size_t j = 1; j < p->r.size(); j++) drmin = std::min(drmin, (double)p->r[j] - p->r[j - 1]);
  if (!(drmin > 0)) drmin = 1e-3;
  const double hw = 0.5 * kTW * std::max(g.dx, g.dy) * 1.41421357 + g.rmargin + 1e-3;
  const double h = p->rl_inv_h > 0 ? 1.0 / p->rl_inv_h : 0.0;
  const int one = (int)std::ceil((2 * hw + 2 * h) / drmin) + 3;
  const int c = p->d_rlut ? 2 * one : one;
  return c < g.nr ? c : g.nr;
}

const char* tile_misfit(const apirl_plan* p) {
  if (p->geo.dims != 3) return "the tile backprojector is 3D only";
  if (p->geo.nsino >= 65536) return "the tile backprojector takes fewer than 65536 sinograms";
  if ((long)p->geo.nang * tile_max_candidates(p) >= 65536)
    return "the tile backprojector takes fewer than 65536 candidate rays per tile";
  if (tile_lds_bytes(p, false) > 65536) return "the tile backprojector's items exceed 64 KiB of LDS";
  return nullptr;
}

int tile_items(const apirl_plan* p) {
  const int zs = tile_slab_depth(p);
  return ((p->geo.nx + kTW - 1) / kTW) * ((p->geo.ny + kTW - 1) / kTW) *
         ((p->geo.nz + zs - 1) / zs);
}

void build_tile_order(const apirl_plan* p, std::vector<int>& order) {
  const int zs = tile_slab_depth(p), nslab = (p->geo.nz + zs - 1) / zs;
  const int ntx = (p->geo.nx + kTW - 1) / kTW, nty = (p->geo.ny + kTW - 1) / kTW;
  std::vector<std::pair<double, int>> tiles;
  for (int ty = 0; ty < nty; ty++)
    for (int tx = 0; tx < ntx; tx++) {
      const double cx = (tx + 0.5) * kTW - 0.5 * p->geo.nx, cy = (ty + 0.5) * kTW - 0.5 * p->geo.ny;
      tiles.push_back({cx * cx + cy * cy, ty * ntx + tx});
    }
  std::stable_sort(tiles.begin(), tiles.end(),
                   [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                     return a.first < b.first;
                   });
  order.clear();
  for (const auto& t : tiles)
    for (int s = 0; s < nslab; s++) order.push_back(t.second * nslab + s);
}

namespace {
size_t align256(size_t n) { return (n + 255) / 256 * 256; }

struct TileSizes {
  int nang_sub, maxc, ntiles;
  long nrays, npair, nslots;
};

TileSizes tile_sizes(const apirl_plan* p, int subset, int nsubsets) {
  const Geo& g = p->geo;
  TileSizes z;
  z.nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  z.npair = (long)z.nang_sub * g.nr;
  z.nrays = (long)g.nsino * z.npair;
  z.maxc = tile_max_candidates(p);
  z.ntiles = ((g.nx + kTW - 1) / kTW) * ((g.ny + kTW - 1) / kTW);
  z.nslots = (long)z.ntiles * z.nang_sub * z.maxc;
  return z;
}
}  // namespace

namespace {
// geometry: [ray count per tile][candidate rays of every (tile, angle)][their paths]
// [per item: entries before each ray, and the total][per item: first entry]
struct TileGeoPtrs {
  int* tcount;
  TileRay* trays;
  TilePath* tpaths;
  int* loff;
  long long* lbase;
  int* tot;
  size_t bytes;
};

TileGeoPtrs tile_geo_ptrs(const apirl_plan* p, const TileSizes& z, void* geo) {
  const int nitems = tile_items(p);
  const int nslot = z.nang_sub * z.maxc;
  TileGeoPtrs q;
  char* c = (char*)geo;
  size_t o = 0;
  q.tcount = (int*)(c + o);
  o += align256((size_t)z.ntiles * sizeof(int));
  q.trays = (TileRay*)(c + o);
  q.tpaths = (TilePath*)(c + o + (size_t)z.nslots * sizeof(TileRay));
  o += align256((size_t)z.nslots * (sizeof(TileRay) + sizeof(TilePath)));
  q.loff = (int*)(c + o);
  o += align256((size_t)nitems * (nslot + 1) * sizeof(int));
  q.lbase = (long long*)(c + o);
  o += align256((size_t)(nitems + 1) * sizeof(long long));
  q.tot = (int*)(c + o);
  o += align256((size_t)nitems * sizeof(int));
  q.bytes = o;
  return q;
}
}  // namespace

size_t tile_geo_bytes(const apirl_plan* p, int subset, int nsubsets) {
  return tile_geo_ptrs(p, tile_sizes(p, subset, nsubsets), nullptr).bytes;
}

hipError_t launch_tile_lists(const apirl_plan* p, int subset, int nsubsets, void* geo,
                             unsigned short* list, long long** total, hipStream_t st) {
  if (tile_misfit(p)) return hipErrorNotSupported;
  const Geo& g = p->geo;
  const TileSizes z = tile_sizes(p, subset, nsubsets);
  const TileGeoPtrs q = tile_geo_ptrs(p, z, geo);
  const int nitems = tile_items(p);
  const int zs = tile_slab_depth(p), nslab = (g.nz + zs - 1) / zs;
  const int nslot = z.nang_sub * z.maxc;
  if (total) *total = q.lbase + nitems;
  GatherArgs A;
  ZRay* unused_zr = nullptr;
  int* unused_nz = nullptr;
  gather_args_for(p, subset, nsubsets, BP_ONES, nullptr, nullptr, nullptr, A, unused_zr,
                  unused_nz);
  const dim3 grid((nitems + kTWaves - 1) / kTWaves);
  const size_t lds = (size_t)g.nsino * sizeof(float2);
  if (!list) {
    hipLaunchKernelGGL(tile_list_kernel<true>, grid, dim3(kGBlock), lds, st, A, zs, nslab, nitems,
                       (const TileRay*)q.trays, (const int*)q.tcount, nslot, q.loff, q.tot,
                       (const long long*)nullptr, (unsigned short*)nullptr);
    hipLaunchKernelGGL(tile_scan_kernel, dim3(1), dim3(kGBlock), 0, st, (const int*)q.tot, nitems,
                       q.lbase);
  } else {
    hipLaunchKernelGGL(tile_list_kernel<false>, grid, dim3(kGBlock), lds, st, A, zs, nslab,
                       nitems, (const TileRay*)q.trays, (const int*)q.tcount, nslot,
                       (int*)nullptr, (int*)nullptr, (const long long*)q.lbase, list);
  }
  return hipGetLastError();
}

// scratch: [ZRay x nrays, (angle, radial bin, sinogram) order][non-zero flag per
// (angle, r), item counter][live positions per slot][live count per tile][the
// geometry, when it is not kept]
size_t tile_scratch_bytes(const apirl_plan* p, int subset, int nsubsets, bool with_geo) {
  const TileSizes z = tile_sizes(p, subset, nsubsets);
  return align256((size_t)z.nrays * sizeof(ZRay)) + align256((size_t)(z.npair + 1) * sizeof(int)) +
         align256((size_t)z.nslots * sizeof(int)) + align256((size_t)z.ntiles * sizeof(int)) +
         (with_geo ? tile_geo_bytes(p, subset, nsubsets) : 0);
}

hipError_t launch_backproject_tile(const apirl_plan* p, const float* sino, float* img,
                                   int subset, int nsubsets, int kind, void* scratch,
                                   size_t scratch_bytes, hipStream_t st, const EmFuse* fuse,
                                   void* geo, bool geo_ready, const unsigned short* list) {
  if (kind != BP_PLAIN && kind != BP_ONES && kind != BP_MULT) return hipErrorNotSupported;
  if (tile_misfit(p)) return hipErrorNotSupported;
  const Geo& g = p->geo;
  const TileSizes z = tile_sizes(p, subset, nsubsets);
  if (z.nrays == 0 || (long)g.nx * g.ny * g.nz == 0) return hipSuccess;
  if (scratch_bytes < tile_scratch_bytes(p, subset, nsubsets, geo == nullptr))
    return hipErrorInvalidValue;
  const int nitems = tile_items(p);
  if (!p->d_tile_order || p->tile_order_n != nitems) return hipErrorInvalidValue;
  char* sp = (char*)scratch;
  ZRay* zr = (ZRay*)sp;
  sp += align256((size_t)z.nrays * sizeof(ZRay));
  int* nzp = (int*)sp;
  sp += align256((size_t)(z.npair + 1) * sizeof(int));
  int* live = (int*)sp;
  sp += align256((size_t)z.nslots * sizeof(int));
  int* nlive = (int*)sp;
  sp += align256((size_t)z.ntiles * sizeof(int));
  if (!geo) {
    geo = sp;
    geo_ready = false;
    list = nullptr;
  }
  const TileGeoPtrs q = tile_geo_ptrs(p, z, geo);
  int* tcount = q.tcount;
  TileRay* trays = q.trays;
  TilePath* tpaths = q.tpaths;
  GatherArgs A;
  ZRay* unused_zr = nullptr;
  int* unused_nz = nullptr;
  gather_args_for(p, subset, nsubsets, kind, sino, img, nullptr, A, unused_zr, unused_nz);
  A.zr = zr;
  A.nz_pair = nzp;
  A.ztab = nullptr;  // the lanes take their z crossings by the loop's additions
  if (fuse) {
    A.upd_sens = fuse->sens;
    A.upd_thr = fuse->thr;
    A.upd_strict = fuse->strict;
  }
  // the call's axial records with their bin values, and the (angle, r) with a
  // non-zero bin
  hipLaunchKernelGGL(clear_int_kernel, dim3(ggrid(z.npair + 1)), dim3(kGBlock), 0, st, nzp,
                     z.npair + 1);
  hipLaunchKernelGGL(tile_zsetup_kernel, dim3((g.nr + 31) / 32, z.nang_sub, (g.nsino + 31) / 32),
                     dim3(kGBlock), 0, st, A, zr, nzp);
  if (!geo_ready) {
    // the subset's geometry: the candidate rays of every (tile, angle) with their
    // paths, compacted per tile
    hipLaunchKernelGGL(tile_rays_kernel, dim3(ggrid(z.nslots)), dim3(kGBlock), 0, st, A, z.maxc,
                       z.nslots, trays, tpaths);
    hipLaunchKernelGGL(tile_compact_kernel, dim3((z.ntiles + kTWaves - 1) / kTWaves),
                       dim3(kGBlock), 0, st, trays, tpaths, z.nang_sub * z.maxc, z.ntiles, tcount);
  }
  // each tile's rays with a non-zero bin in this call
  hipLaunchKernelGGL(tile_live_kernel, dim3((z.ntiles + kTWaves - 1) / kTWaves), dim3(kGBlock), 0,
                     st, (const TileRay*)trays, (const int*)tcount, (const int*)nzp,
                     z.nang_sub * z.maxc, z.ntiles, live, nlive);
  const int zs = tile_slab_depth(p);
  const int nslab = (g.nz + zs - 1) / zs;
  const dim3 grid((unsigned)((nitems + kTWaves - 1) / kTWaves));
  if (list && geo_ready)
    hipLaunchKernelGGL(tile_bp_kernel<true>, grid, dim3(kGBlock), tile_lds_bytes(p, true), st, A, zs,
                       nslab, p->d_tile_order, nitems, nzp + z.npair, (const TileRay*)trays,
                       (const TilePath*)tpaths, z.nang_sub * z.maxc, (const int*)live,
                       (const int*)nlive, (const int*)q.loff, (const long long*)q.lbase, list);
  else
    hipLaunchKernelGGL(tile_bp_kernel<false>, grid, dim3(kGBlock), tile_lds_bytes(p, false), st, A, zs,
                       nslab, p->d_tile_order, nitems, nzp + z.npair, (const TileRay*)trays,
                       (const TilePath*)tpaths, z.nang_sub * z.maxc, (const int*)live,
                       (const int*)nlive, (const int*)nullptr, (const long long*)nullptr,
                       (const unsigned short*)nullptr);
  return hipGetLastError();
}

// Bytes of the z-crossing table for a subset of nang_sub angles; 0 (no table: the
// gather reconstructs z crossings with alpha_after) above 1 GiB, e.g. for MLEM over
// all angles of a large sinogram.
size_t ztable_bytes(const apirl_plan* p, int nang_sub) {
  const size_t bytes = (size_t)nang_sub * p->geo.nr * (size_t)p->zblock * sizeof(float);
  return (p->d_zsoff && bytes <= ((size_t)1 << 30)) ? bytes : 0;
}

void gather_args_for(const apirl_plan* p, int subset, int nsubsets, int kind, const float* sino,
                     float* img, void* scratch, GatherArgs& A, ZRay*& zr, int*& nz_pair) {
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  A.g = g;
  A.rays = p->d_rays;
  A.zab = p->d_zab;
  A.angcs = p->d_angcs;
  A.segs = p->d_segs;
  A.nseg = p->nseg;
  A.sino = sino;
  A.img = img;
  A.kind = kind;
  A.nang_sub = nang_sub;
  A.subset = subset;
  A.nsubsets = nsubsets;
  A.zlanes = 1;
  zr = nullptr;
  nz_pair = nullptr;
  A.ztab = nullptr;
  A.zsoff = p->d_zsoff;
  A.zblock = p->zblock;
  A.xytab = p->d_xytab;
  A.angcs_odd = p->d_angcs_odd;
  A.rlut = p->d_rlut;
  A.rl0 = p->rl0;
  A.rl_inv_h = p->rl_inv_h;
  A.nrl = p->nrl;
  // the column kernel's sector order when selected and built for the current depth
  const int zv = p->gather_zvox;
  const bool wedge = p->gather_xcd == APIRL_GATHER_XCD_WEDGES && zv >= 1 && zv <= 4 &&
                     p->d_tilemap[zv];
  A.tile_map = wedge ? p->d_tilemap[zv] : nullptr;
  A.tile_slots = wedge ? p->tilemap_slots[zv] : 0;
  if (g.dims == 3 && scratch) {
    // scratch: [ZRay x nrays][non-zero flag per (angle, r)][z-crossing table]
    const size_t npair = (size_t)nang_sub * g.nr;
    zr = (ZRay*)scratch;
    nz_pair = (int*)((char*)scratch + (size_t)nrays * sizeof(ZRay));
    if (ztable_bytes(p, nang_sub) > 0)
      A.ztab = (float*)((char*)scratch + (size_t)nrays * sizeof(ZRay) +
                        (npair * sizeof(int) + 255) / 256 * 256);
  }
  A.zr = zr;
  A.nz_pair = nz_pair;
  A.upd_sens = nullptr;
  A.upd_thr = 0.f;
  A.upd_strict = 0;
}

hipError_t launch_group_gather(BatchHeader* hdr, const ZsJob* zs, const GJob* g2,
                               const GJob* g3, int grid, int max_nsino, hipStream_t st) {
  hipLaunchKernelGGL(zsetup_group_kernel, dim3(grid), dim3(kGBlock), 0, st, hdr, zs);
  hipLaunchKernelGGL(gather_group_kernel<2>, dim3(grid), dim3(kGBlock), 0, st, hdr, g2);
  hipLaunchKernelGGL(gather_group_kernel<3>, dim3(grid), dim3(kGBlock),
                     column_lds_bytes(max_nsino), st, hdr, g3);
  return hipGetLastError();
}

const char* gather_misfit(const apirl_plan* p, int subset, int nsubsets) {
  const Geo& g = p->geo;
  if (g.dims != 3) return nullptr;
  const long nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long npair = nang_sub * g.nr;
  const long lim24 = 1L << 24;
  if ((long)g.nsino * npair >= (1L << 31)) return "the subset has 2^31 bins or more";
  if (npair >= lim24) return "the subset has 2^24 (angle, radial bin) pairs or more";
  if (g.nsino >= lim24) return "the sinogram count reaches 2^24";
  if (p->zblock >= lim24) return "the z-crossing table slots reach 2^24 entries";
  // the table offset e * npair (e < zblock) is a 32-bit product; the table exists only
  // below 1 GiB, 2^28 floats, which keeps it there, and this says so explicitly
  if (ztable_bytes(p, (int)nang_sub) > 0 && (unsigned long)p->zblock * npair >= (1UL << 32))
    return "the z-crossing table offsets reach 2^32";
  return nullptr;
}

void build_wedge_tile_map(int ntx, int nty, int nzb, std::vector<int>& map) {
  constexpr int kX = 8;  // XCDs; the hardware deals workgroup b to XCD b % 8
  std::vector<int> wedge((size_t)ntx * nty);
  const double kPi = 3.14159265358979323846;
  for (int ty = 0; ty < nty; ty++)
    for (int tx = 0; tx < ntx; tx++) {
      const double cx = tx + 0.5 - ntx / 2.0, cy = ty + 0.5 - nty / 2.0;
      int w = (int)std::floor((std::atan2(cy, cx) + kPi) / (2 * kPi / kX));
      wedge[(size_t)ty * ntx + tx] = w < 0 ? 0 : (w >= kX ? kX - 1 : w);
    }
  std::vector<std::vector<int>> lists(kX);
  for (int tz = 0; tz < nzb; tz++)
    for (int ty = 0; ty < nty; ty++)
      for (int tx = 0; tx < ntx; tx++)
        lists[wedge[(size_t)ty * ntx + tx]].push_back(tx + ntx * (ty + nty * tz));
  size_t len = 0;
  for (auto& l : lists) len = l.size() > len ? l.size() : len;
  map.assign(len * kX, -1);
  for (int x = 0; x < kX; x++)
    for (size_t k = 0; k < lists[x].size(); k++) map[k * kX + x] = lists[x][k];
}

bool gather_fits(const apirl_plan* p, int subset, int nsubsets) {
  return gather_misfit(p, subset, nsubsets) == nullptr;
}

size_t gather_scratch_bytes(const apirl_plan* p, int subset, int nsubsets) {
  const Geo& g = p->geo;
  if (g.dims != 3) return 0;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  const size_t npair = (size_t)nang_sub * g.nr;
  return (size_t)nrays * sizeof(ZRay) + (npair * sizeof(int) + 255) / 256 * 256 +
         ztable_bytes(p, nang_sub);
}

}  // namespace apirl
