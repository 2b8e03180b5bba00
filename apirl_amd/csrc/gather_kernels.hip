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
#include <cmath>

#include "apirl_internal.h"
#include "alpha_exact.h"

namespace apirl {

namespace {

constexpr int kGBlock = 256;
// axial margin of the sinogram search: covers the drift of the accumulated z
// crossings against the exact slab faces (well below 1e-3 mm here)
constexpr float kZMargin = 0.05f;

__device__ __forceinline__ float z_at(float za, float vz, float alpha) { return za + vz * alpha; }

// Per-ray axial part of the Siddon setup for one subset (Siddon.cpp:150-182,
// 220, 244-255, 279-286) and the bin value, packed in one 32-byte record so a
// candidate ray costs one load: computed once per backprojection by zsetup_kernel.
struct alignas(32) ZRay {
  int k0;       // entry slice
  short k1;     // exit slice
  short kinc;   // sign of Vz
  float azu;    // |dz / Vz|
  float az0;    // first z crossing (FLT_MAX when Vz == 0)
  float L;      // ray length |P2 - P1| in the reference's operand order
  float b;      // bin value being backprojected
  float pad0, pad1;
};
static_assert(sizeof(ZRay) == 32, "ZRay is one 32-byte load");

// nz_pair[ka * nr + r] != 0 iff some sinogram has a non-zero bin at (angle ka, r):
// the gather skips (angle, r) pairs whose bins are all zero (zero bins add nothing,
// SiddonProjector.cpp:496).
__global__ __launch_bounds__(kGBlock) void zsetup_kernel(Geo g, const RayXY* __restrict__ rays,
                                                         const float2* __restrict__ zab,
                                                         const float* __restrict__ sino, int kind,
                                                         int nang_sub, int subset, int nsubsets,
                                                         long nrays, ZRay* __restrict__ out,
                                                         int* __restrict__ nz_pair) {
  const long idx = (long)blockIdx.x * kGBlock + threadIdx.x;
  if (idx >= nrays) return;
  const int rr = (int)(idx % g.nr);
  const long t = idx / g.nr;
  const int ka = (int)(t % nang_sub);
  const int s = (int)(t / nang_sub);
  const int a = subset + nsubsets * ka;
  const RayXY R = rays[(long)a * g.nr + rr];
  const float za = zab[s].x, zb = zab[s].y;
  const float vz = zb - za;
  const float z1 = za + vz * R.amin;
  const float z2 = za + vz * R.amax;
  ZRay z;
  z.k0 = (int)floorf((z1 - g.offz) / g.dz);
  z.k1 = (short)(int)floorf((z2 - g.offz) / g.dz);
  z.azu = fabsf(g.dz / vz);
  const int kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
  z.kinc = (short)kinc;
  if (vz > 0) z.az0 = (g.offz + (z.k0 + kinc) * g.dz - za) / vz;
  else if (vz < 0) z.az0 = (g.offz + z.k0 * g.dz - za) / vz;
  else z.az0 = FLT_MAX;
  const float ez = (za + vz) - za;
  z.L = sqrtf(R.lxy + ez * ez);
  float b = 1.f;
  if (kind == BP_PLAIN) b = sino[idx];
  else if (kind == BP_MULT) b = sino[((long)s * g.nang + a) * g.nr + rr];
  z.b = b;
  z.pad0 = z.pad1 = 0.f;
  out[idx] = z;
  if (b != 0.f && (R.flags & RAY_VALID)) {
    int* f = &nz_pair[(long)ka * g.nr + rr];
    if (*f == 0) atomicOr(f, 1);
  }
}

__global__ __launch_bounds__(kGBlock) void clear_int_kernel(int* __restrict__ p, long n) {
  const long i = (long)blockIdx.x * kGBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

struct GatherArgs {
  Geo g;
  const RayXY* rays;
  const float2* zab;
  const float2* angcs;  // (cos, sin) of every angle of the full sinogram
  const GatherSeg* segs;
  int nseg;
  const ZRay* zr;
  const int* nz_pair;
  const float* sino;  // subset sinogram (PLAIN), full sinogram (MULT) or null (ONES)
  float* img;
  int kind;           // BP_PLAIN, BP_ONES or BP_MULT
  int nang_sub, subset, nsubsets;
};

// Voxel tile of one workgroup: 16 x 16 voxels of one slice, so the lanes of a wave
// (4 rows of 16) share most candidate rays and sinograms.
constexpr int kTile = 16;

template <int DIMS>
__global__ __launch_bounds__(kGBlock) void gather_kernel(GatherArgs A) {
  extern __shared__ float2 s_zab[];  // (P1.Z, P2.Z) of every sinogram
  const Geo& g = A.g;
  if (DIMS == 3) {
    for (int q = threadIdx.x; q < g.nsino; q += kGBlock) s_zab[q] = A.zab[q];
    __syncthreads();
  }
  const int i = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int j = blockIdx.y * kTile + (threadIdx.x / kTile);
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const long v = ((long)k * g.ny + j) * g.nx + i;
  // voxel centre for the candidate search
  const float xc = -g.R + g.dx * ((float)i + 0.5f);
  const float yc = -g.R + g.dy * ((float)j + 0.5f);
  const float zk0 = g.offz + k * g.dz, zk1 = zk0 + g.dz;
  const float inv_dr = 1.0f / g.dr;
  float acc = A.img[v];
  for (int ka = 0; ka < A.nang_sub; ka++) {
    const int a = A.subset + A.nsubsets * ka;
    const float2 cs = A.angcs[a];
    const float sc = xc * cs.x + yc * cs.y;
    const float hw = 0.5f * g.dx * (fabsf(cs.x) + fabsf(cs.y)) + g.rmargin;
    int jlo = (int)ceilf((sc - hw - g.r0) * inv_dr);
    int jhi = (int)floorf((sc + hw - g.r0) * inv_dr);
    jlo = jlo < 0 ? 0 : jlo;
    jhi = jhi > g.nr - 1 ? g.nr - 1 : jhi;
    for (int rr = jlo; rr <= jhi; rr++) {
      if (DIMS == 3 && A.nz_pair[(long)ka * g.nr + rr] == 0) continue;
      const RayXY t = A.rays[(long)a * g.nr + rr];
      if (!(t.flags & RAY_VALID)) continue;
      const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
      int na, nb;
      if (iinc == 0) {
        if (i != t.i0) continue;
        na = 0;
      } else {
        na = (i - t.i0) * iinc;
        if (na < 0) continue;
      }
      if (jinc == 0) {
        if (j != t.j0) continue;
        nb = 0;
      } else {
        nb = (j - t.j0) * jinc;
        if (nb < 0) continue;
      }
      if (na + nb >= t.nxy + (DIMS == 3 ? g.nz + 1 : 0)) continue;
      // x / y crossings around the column
      float Xp = -FLT_MAX, Xn = t.ax, Yp = -FLT_MAX, Yn = t.ay;
      if (na > 0) {
        Xp = alpha_after(t.ax, t.axu, na - 1);
        Xn = Xp + t.axu;
      }
      if (nb > 0) {
        Yp = alpha_after(t.ay, t.ayu, nb - 1);
        Yn = Yp + t.ayu;
      }
      // taken x before untaken y (x wins ties), taken y before untaken x (strict)
      if (na > 0 && !(Xp <= Yn)) continue;
      if (nb > 0 && !(Yp < Xn)) continue;
      if (DIMS == 2) {
        const int m = na + nb;
        if (m >= t.nxy) continue;
        float ac = t.amin;
        if (m > 0) ac = fmaxf(Xp, Yp);
        const float an = fminf(Xn, Yn);
        const float w = (an - ac) * sqrtf(t.lxy);
        float b = 1.f;
        if (A.kind == BP_PLAIN) b = A.sino[(long)ka * g.nr + rr];
        else if (A.kind == BP_MULT) b = A.sino[(long)a * g.nr + rr];
        acc += w * b * 1.0f;  // SiddonProjector.cpp:72
        continue;
      }
      // alpha interval of the column; the sinograms whose ray crosses the slab
      // [zk0, zk1] inside it form a contiguous run of each segment (the mean ring
      // coordinates grow with the index), found by a short scan from an estimate.
      const float lo = fmaxf(t.amin, fmaxf(Xp, Yp));
      const float hi = fminf(t.amax, fminf(Xn, Yn));
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
          const long ridx = ((long)s * A.nang_sub + ka) * g.nr + rr;
          const ZRay z = A.zr[ridx];
          if (z.b == 0.f) continue;  // SiddonProjector.cpp:496
          int nc;
          if (z.kinc == 0) {
            if (k != z.k0) continue;
            nc = 0;
          } else {
            nc = (k - z.k0) * z.kinc;
            if (nc < 0) continue;
          }
          const int m = na + nb + nc;
          if (m >= t.nxy + abs((int)z.k1 - z.k0)) continue;
          float Zp = -FLT_MAX, Zn = z.az0;
          if (nc > 0) {
            Zp = alpha_after(z.az0, z.azu, nc - 1);
            Zn = Zp + z.azu;
          }
          // cross-axis order (x before y before z on ties)
          if (na > 0 && !(Xp <= Zn)) continue;
          if (nb > 0 && !(Yp <= Zn)) continue;
          if (nc > 0 && !(Zp < Xn && Zp < Yn)) continue;
          float ac = t.amin;
          if (m > 0) ac = fmaxf(fmaxf(Xp, Yp), Zp);
          const float an = fminf(fminf(Xn, Yn), Zn);
          const float w = (an - ac) * z.L;
          acc += w * 1.0f * z.b;  // SiddonProjector.cpp:546
        }
      }
    }
  }
  A.img[v] = acc;
}

inline unsigned ggrid(long n) { return (unsigned)((n + kGBlock - 1) / kGBlock); }

}  // namespace

hipError_t launch_backproject_gather(const apirl_plan* p, const float* sino, float* img,
                                     int subset, int nsubsets, int kind, void* scratch,
                                     size_t scratch_bytes, hipStream_t st) {
  if (kind != BP_PLAIN && kind != BP_ONES && kind != BP_MULT) return hipErrorNotSupported;
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  const long nvox = (long)g.nx * g.ny * g.nz;
  if (nrays == 0 || nvox == 0) return hipSuccess;
  GatherArgs A;
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
  A.zr = nullptr;
  A.nz_pair = nullptr;
  if (g.dims == 3) {
    const size_t need = gather_scratch_bytes(p, subset, nsubsets);
    if (scratch_bytes < need) return hipErrorInvalidValue;
    ZRay* zr = (ZRay*)scratch;
    int* nzp = (int*)((char*)scratch + (size_t)nrays * sizeof(ZRay));
    const long npair = (long)nang_sub * g.nr;
    hipLaunchKernelGGL(clear_int_kernel, dim3(ggrid(npair)), dim3(kGBlock), 0, st, nzp, npair);
    hipLaunchKernelGGL(zsetup_kernel, dim3(ggrid(nrays)), dim3(kGBlock), 0, st, g, p->d_rays,
                       p->d_zab, sino, kind, nang_sub, subset, nsubsets, nrays, zr, nzp);
    A.zr = zr;
    A.nz_pair = nzp;
    const dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, g.nz);
    hipLaunchKernelGGL(gather_kernel<3>, grid, dim3(kGBlock), (size_t)g.nsino * sizeof(float2),
                       st, A);
  } else {
    const dim3 grid((g.nx + kTile - 1) / kTile, (g.ny + kTile - 1) / kTile, g.nz);
    hipLaunchKernelGGL(gather_kernel<2>, grid, dim3(kGBlock), 0, st, A);
  }
  return hipGetLastError();
}

size_t gather_scratch_bytes(const apirl_plan* p, int subset, int nsubsets) {
  const Geo& g = p->geo;
  if (g.dims != 3) return 0;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  return (size_t)nrays * sizeof(ZRay) + (size_t)nang_sub * g.nr * sizeof(int);
}

}  // namespace apirl
