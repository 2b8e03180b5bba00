// gfx950 kernels of the Siddon projector / backprojector and the EM update.
//
// One ray per lane.  Lanes of a wavefront take consecutive radial bins of one
// (sinogram, angle), so the 64-byte RayXY records, the measured / additive bins and
// the sinogram writes are contiguous per wavefront, and the 64 rays are parallel
// neighbours whose voxel gathers hit the same cache lines of the (L2 / MALL
// resident) image.  The path is gather/scatter bound; no MFMA is involved.
//
// Arithmetic follows recon/src/Siddon.cpp in float, operation by operation, and the
// file is compiled with -ffp-contract=off and correctly rounded division / sqrt, so
// every voxel index and weight equals the reference's.  Accumulation along a ray
// is sequential in traversal order, which is the reference's order
// (SiddonProjector.cpp:176-181, 413-423): the forward projection is bit-exact.
#include <cfloat>
#include <cmath>

#include "apirl_internal.h"
#include "siddon_f64.h"

namespace apirl {

namespace {

constexpr int kBlock = 256;
constexpr int kNumXcd = 8;  // MI355X: 8 XCDs of 32 CUs
constexpr int kUnroll = 8;  // Siddon steps per batch of voxel loads
constexpr int kPathUnroll = 12;  // the same for project_paths_kernel (12: 0.86 ms per C2 subset; 8: 0.90, 16: 0.88)
// Marks a batch slot without a term: no Siddon weight is NaN-boxed like this value.

// Per-ray traversal state (Siddon.cpp:253-340).
struct Ray {
  float ax, ay, az, axu, ayu, azu, ac, L;
  int i, j, k, iinc, jinc, kinc, n;
};

// Builds the traversal state of bin (s, angle a, r) from the transaxial table and
// the sinogram's axial coordinates (Siddon.cpp:150-182, 220, 234-255, 279-286).
template <int DIMS>
__device__ __forceinline__ bool setup_ray(const Geo& g, const RayXY& t, float za, float zb,
                                          Ray& ray) {
  // Every field is assigned unconditionally at the end from locals, so the compiler
  // keeps the state in registers (no address-selected stores into a private struct).
  float az = FLT_MAX, azu = 0.f, L;
  int k = 0, kinc = 0, n = t.nxy;
  if (DIMS == 2) {
    L = sqrtf(t.lxy);
  } else {
    const float vz = zb - za;
    const float z1 = za + vz * t.amin;
    const float z2 = za + vz * t.amax;
    const int k0 = (int)floorf((z1 - g.offz) / g.dz);
    const int k1 = (int)floorf((z2 - g.offz) / g.dz);
    k = k0;
    kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
    n = t.nxy + abs(k1 - k0);
    const float ez = (za + vz) - za;
    L = sqrtf(t.lxy + ez * ez);
    azu = fabsf(g.dz / vz);
    if (vz > 0) az = (g.offz + (k0 + kinc) * g.dz - za) / vz;
    else if (vz < 0) az = (g.offz + k0 * g.dz - za) / vz;
  }
  ray.ax = t.ax;
  ray.ay = t.ay;
  ray.az = az;
  ray.axu = t.axu;
  ray.ayu = t.ayu;
  ray.azu = azu;
  ray.ac = t.amin;
  ray.L = L;
  ray.i = t.i0;
  ray.j = t.j0;
  ray.k = k;
  ray.iinc = ray_iinc(t.flags);
  ray.jinc = ray_jinc(t.flags);
  ray.kinc = kinc;
  ray.n = n;
  return (t.flags & RAY_VALID) != 0;
}

// One Siddon step: returns the weight of the current voxel (ci, cj, ck) and advances
// (Siddon.cpp:310-340 for 3D, :615-639 for 2D; the 2D form compares only x and y,
// which is the 3D rule with az = FLT_MAX).
__device__ __forceinline__ float step(Ray& r, int& ci, int& cj, int& ck) {
  ci = r.i;
  cj = r.j;
  ck = r.k;
  const bool bx = (r.ax <= r.ay) && (r.ax <= r.az);
  const bool by = !bx && (r.ay <= r.az);
  const float an = bx ? r.ax : (by ? r.ay : r.az);
  const float w = (an - r.ac) * r.L;
  r.ac = an;
  if (bx) {
    r.i += r.iinc;
    r.ax += r.axu;
  } else if (by) {
    r.j += r.jinc;
    r.ay += r.ayu;
  } else {
    r.k += r.kinc;
    r.az += r.azu;
  }
  return w;
}

// Branch-free form of the same step for the hot loops: the three candidate updates
// are computed and selected (a taken alpha advances by exactly the reference's
// addition; an untaken one keeps its value), and the voxel's linear index li is
// carried along (+-1 in x, +-nx in y, +-nx*ny in z) instead of being recomputed.
struct RayIdx {
  int ci, cj, ck, li;  // voxel emitted by the step and its linear index
};

__device__ __forceinline__ float step_bf(Ray& r, RayIdx& v, int& li, int jstep, int kstep) {
  const bool bx = (r.ax <= r.ay) && (r.ax <= r.az);
  const bool by = !bx && (r.ay <= r.az);
  const bool bz = !bx && !by;
  const float an = bx ? r.ax : (by ? r.ay : r.az);
  const float w = (an - r.ac) * r.L;
  r.ac = an;
  const float nax = r.ax + r.axu, nay = r.ay + r.ayu, naz = r.az + r.azu;
  r.ax = bx ? nax : r.ax;
  r.ay = by ? nay : r.ay;
  r.az = bz ? naz : r.az;
  v.ci = r.i;
  v.cj = r.j;
  v.ck = r.k;
  v.li = li;
  r.i += bx ? r.iinc : 0;
  r.j += by ? r.jinc : 0;
  r.k += bz ? r.kinc : 0;
  li += bx ? r.iinc : (by ? jstep : kstep);
  return w;
}

// A ray's sum in traversal order: the reference's float additions, or (KAHAN) the
// same terms with a Kahan compensation term (APIRL_PREC_KAHAN, APIRL_PREC_F64_SETUP).
template <bool KAHAN>
struct RaySum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float term) {
    if (KAHAN) {
      const float y = term - c;
      const float t = s + y;
      c = (t - s) - y;
      s = t;
    } else {
      s = s + term;
    }
  }
  __device__ __forceinline__ void add_if(bool ok, float term) {
    RaySum n = *this;
    n.add(term);
    s = ok ? n.s : s;
    c = ok ? n.c : c;
  }
};

// kUnroll steps of walk3d (below), from step m0, with the read test on every step;
// the batch may run past the ray's last step, whose terms are then not added.
// STRICT: a voxel is read when it lies inside the image (the F64_SETUP loop's test);
// otherwise the reference's test (upper bounds and li >= 0).
template <bool TR, bool KAHAN, bool STRICT>
__device__ __forceinline__ void walk3d_batch(const Geo& g, Ray& ray, const float* __restrict__ img,
                                             const float* __restrict__ imgT, int nxy, int m0,
                                             RaySum<KAHAN>& acc) {
  float wv[kUnroll], pv[kUnroll];
  bool tv[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; u++) {
    const int ci = ray.i, cj = ray.j, ck = ray.k;
    const bool bx = (ray.ax <= ray.ay) && (ray.ax <= ray.az);
    const bool by = !bx && (ray.ay <= ray.az);
    const bool bz = !bx && !by;
    const float an = bx ? ray.ax : (by ? ray.ay : ray.az);
    wv[u] = (an - ray.ac) * ray.L;
    ray.ac = an;
    const float nax = ray.ax + ray.axu, nay = ray.ay + ray.ayu, naz = ray.az + ray.azu;
    ray.ax = bx ? nax : ray.ax;
    ray.ay = by ? nay : ray.ay;
    ray.az = bz ? naz : ray.az;
    ray.i += bx ? ray.iinc : 0;
    ray.j += by ? ray.jinc : 0;
    ray.k += bz ? ray.kinc : 0;
    const int li = __mul24(ck, nxy) + __mul24(cj, g.nx) + ci;
    const bool ok = STRICT ? ((unsigned)ci < (unsigned)g.nx) & ((unsigned)cj < (unsigned)g.ny) &
                                 ((unsigned)ck < (unsigned)g.nz) & (m0 + u < ray.n)
                           : (ci < g.nx) & (cj < g.ny) & (ck < g.nz) & (li >= 0) &
                                 (m0 + u < ray.n);
    int idx = li;
    const float* base = img;
    if (TR && ((ci >= 0) & (cj >= 0))) {
      idx = __mul24(ck, nxy) + __mul24(ci, g.ny) + cj;
      base = imgT;
    }
    tv[u] = ok;
    // idx >= 0 whenever ok holds, so the offset is taken unsigned (no sign extension)
    pv[u] = base[(unsigned)(ok ? idx : 0)];
  }
#pragma unroll
  for (int u = 0; u < kUnroll; u++) acc.add_if(tv[u], wv[u] * pv[u]);
}

// kUnroll steps of walk3d from step m0 for a batch that stays inside the image.  A
// ray's voxel index moves monotonically along each axis, so when the first and the
// last voxel of a batch lie inside the image every voxel between them does: the
// read tests all pass and each term reads the in-image voxel, so they and the
// out-of-image index are skipped.  A lane whose batch reaches outside the image, or
// past the ray's last step, takes walk3d_batch from the same state.
template <bool TR, bool KAHAN, bool STRICT>
__device__ __forceinline__ void walk3d_batch_inside(const Geo& g, Ray& ray,
                                                    const float* __restrict__ img,
                                                    const float* __restrict__ imgT, int nxy,
                                                    int m0, RaySum<KAHAN>& acc) {
  Ray r = ray;
  float wv[kUnroll];
  int iv[kUnroll];
  int ei = 0, ej = 0, ek = 0;  // the batch's last voxel
#pragma unroll
  for (int u = 0; u < kUnroll; u++) {
    const int ci = r.i, cj = r.j, ck = r.k;
    // the step's axis, from three comparisons (the tested form's rule: x, then y, then z)
    const bool le_xy = r.ax <= r.ay, le_xz = r.ax <= r.az, le_yz = r.ay <= r.az;
    const bool bx = le_xy & le_xz;
    const bool by = !bx & le_yz;
    const bool bz = !bx & !le_yz;
    const float an = bx ? r.ax : (le_yz ? r.ay : r.az);
    wv[u] = (an - r.ac) * r.L;
    r.ac = an;
    const float nax = r.ax + r.axu, nay = r.ay + r.ayu, naz = r.az + r.azu;
    r.ax = bx ? nax : r.ax;
    r.ay = by ? nay : r.ay;
    r.az = bz ? naz : r.az;
    r.i += bx ? r.iinc : 0;
    r.j += by ? r.jinc : 0;
    r.k += bz ? r.kinc : 0;
    iv[u] = TR ? __mul24(ck, nxy) + __mul24(ci, g.ny) + cj
               : __mul24(ck, nxy) + __mul24(cj, g.nx) + ci;
    if (u == kUnroll - 1) {
      ei = ci;
      ej = cj;
      ek = ck;
    }
  }
  const bool inside = ((unsigned)ray.i < (unsigned)g.nx) & ((unsigned)ray.j < (unsigned)g.ny) &
                      ((unsigned)ray.k < (unsigned)g.nz) & ((unsigned)ei < (unsigned)g.nx) &
                      ((unsigned)ej < (unsigned)g.ny) & ((unsigned)ek < (unsigned)g.nz) &
                      (m0 + kUnroll <= ray.n);
  if (inside) {
    const float* __restrict__ base = TR ? imgT : img;
    float pv[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) pv[u] = base[(unsigned)iv[u]];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) acc.add(wv[u] * pv[u]);
    ray = r;
  } else {
    walk3d_batch<TR, KAHAN, STRICT>(g, ray, img, imgT, nxy, m0, acc);
  }
}

// 3D traversal of one ray, kUnroll steps per batch of voxel loads (addresses never
// depend on loaded data), terms added in traversal order.  The reference reads voxel
// (ci, cj, ck) when ci < nx, cj < ny, ck < nz and the linear index li = ck nx ny +
// cj nx + ci is not negative (SiddonProjector.cpp:417-423; a negative ci or cj then
// reads a neighbouring voxel of the array, as there).  li <= nx ny nz - 1 follows
// from the three upper bounds.  TR: the rays are offset along y, so in-image voxels
// are read from the transposed copy (y fastest), where the wavefront's 64 parallel
// rays touch neighbouring addresses; out-of-image reads keep the reference's index
// into the image.  Indices use 24-bit multiplies (the plan keeps nx ny below 2^23).
// Batches inside the image skip the bounds and step-count tests; the others, and the
// last, partial batch, take them.  KAHAN / STRICT: see RaySum and walk3d_batch.
template <bool TR, bool KAHAN = false, bool STRICT = false>
__device__ __forceinline__ void walk3d(const Geo& g, Ray& ray, const float* __restrict__ img,
                                       const float* __restrict__ imgT, RaySum<KAHAN>& acc) {
  const int nxy = g.nx * g.ny;
  int m0 = 0;
  for (; m0 + kUnroll <= ray.n; m0 += kUnroll)
    walk3d_batch_inside<TR, KAHAN, STRICT>(g, ray, img, imgT, nxy, m0, acc);
  if (m0 < ray.n) walk3d_batch<TR, KAHAN, STRICT>(g, ray, img, imgT, nxy, m0, acc);
}

// Decomposes a ray index of a subset sinogram into (sinogram, angle, r).
__device__ __forceinline__ void ray_coords(const Geo& g, long idx, int nang_sub, int subset,
                                           int nsubsets, int& s, int& a, int& r, int& k) {
  r = (int)(idx % g.nr);
  long t = idx / g.nr;
  k = (int)(t % nang_sub);
  s = (int)(t / nang_sub);
  a = subset + nsubsets * k;
}

// One ray of a (subset) projection: lidx is the ray's position in the z-sorted
// order of the subset's rays.  Returns this lane's likelihood term (fused mode).
template <int PREC, bool FUSED, int DIMS>
__device__ __forceinline__ double project_ray(
    const Geo& g, const RayXY* __restrict__ rays, const float2* __restrict__ zab,
    const float* __restrict__ img, float* __restrict__ sino, int nang_sub, int subset,
    int nsubsets, long nrays, const FusedRatio& fr, const int* __restrict__ zorder,
    const float* __restrict__ imgT, long lidx) {
  double lik = 0.0;
  long idx = 0;
  if (lidx < nrays) {
    int s, a, rr, k;
    ray_coords(g, lidx, nang_sub, subset, nsubsets, s, a, rr, k);
    if (zorder) s = zorder[s];
    idx = ((long)s * nang_sub + k) * g.nr + rr;
    // tables hold ns (transaxial) x nas (axial) LORs per bin; one each without
    // oversampling
    const long rbase = ((long)a * g.nr + rr) * g.ns;
    const RayXY t = rays[rbase];
    float za = 0.f, zb = 0.f;
    if (DIMS == 3) {
      const float2 z = zab[(long)s * g.nas];
      za = z.x;
      zb = z.y;
    }
    Ray ray;
    bool valid = setup_ray<DIMS>(g, t, za, zb, ray);
    float acc = 0.f;
    if (DIMS == 3 && PREC == APIRL_PREC_REFERENCE && g.walk24 && g.ns * g.nas > 1) {
      // oversampled bin: its LORs (o, p) in the reference's order, every term added to
      // the one bin total (SiddonProjector.cpp:366-425)
      valid = false;
      RaySum<false> sum;
      for (int o = 0; o < g.ns; o++) {
        const RayXY to = rays[rbase + o];
        for (int q = 0; q < g.nas; q++) {
          const float2 z = zab[(long)s * g.nas + q];
          Ray ro;
          if (!setup_ray<DIMS>(g, to, z.x, z.y, ro)) continue;
          valid = true;
          if (imgT != nullptr && (to.flags & RAY_XMAJOR)) walk3d<true>(g, ro, img, imgT, sum);
          else walk3d<false>(g, ro, img, imgT, sum);
        }
      }
      acc = sum.s;
    } else if (valid) {
      const bool tr3 = imgT != nullptr && (t.flags & RAY_XMAJOR);
      if (PREC == APIRL_PREC_REFERENCE && DIMS == 3 && g.walk24) {
        RaySum<false> sum;
        if (tr3) walk3d<true>(g, ray, img, imgT, sum);
        else walk3d<false>(g, ray, img, imgT, sum);
        acc = sum.s;
      } else if (PREC == APIRL_PREC_KAHAN && DIMS == 3 && g.walk24) {
        // the reference's traversal with a Kahan-compensated sum
        RaySum<true> sum;
        if (tr3) walk3d<true, true>(g, ray, img, imgT, sum);
        else walk3d<false, true>(g, ray, img, imgT, sum);
        acc = sum.s;
      } else if (PREC == APIRL_PREC_F64_SETUP && DIMS == 3 && g.walk24) {
        // F64_SETUP: the setup in double from the float LOR (siddon_f64.h), rounded to
        // float; then the reference's float loop (Siddon.cpp:310-340: repeated
        // additions, the x / y / z tie order) over in-image voxels with a
        // Kahan-compensated float sum
        SetupXY64 xy;
        SetupZ64 zs;
        setup_xy_f64(g, t.p0x, t.p0y, t.vx, t.vy, xy);
        setup_z_f64(g, xy, za, t.vx, t.vy, (double)(zb - za), zs);
        Ray r64;
        r64.ax = (float)xy.ax;
        r64.ay = (float)xy.ay;
        r64.az = (float)zs.az;
        r64.axu = (float)xy.axu;
        r64.ayu = (float)xy.ayu;
        r64.azu = (float)zs.azu;
        r64.ac = (float)xy.amin;
        r64.L = (float)zs.L;
        r64.i = xy.i0;
        r64.j = xy.j0;
        r64.k = zs.k0;
        r64.iinc = xy.iinc;
        r64.jinc = xy.jinc;
        r64.kinc = zs.kinc;
        r64.n = abs(xy.i1 - xy.i0) + abs(xy.j1 - xy.j0) + abs(zs.k1 - zs.k0);
        RaySum<true> sum;
        if (tr3) walk3d<true, true, true>(g, r64, img, imgT, sum);
        else walk3d<false, true, true>(g, r64, img, imgT, sum);
        acc = sum.s;
      } else if (PREC == APIRL_PREC_REFERENCE) {
        // Steps are taken kUnroll at a time: the voxel loads of a batch are issued
        // together (addresses never depend on loaded data; a step outside the image
        // loads voxel 0 and is not added), then added in traversal order, so the sum
        // is still the reference's sequential one.
        // Voxel reads come from a layout where the 64 parallel rays of a wavefront
        // (neighbouring radial bins) touch neighbouring addresses: the image itself
        // (x fastest) when they are offset along x, its transpose (y fastest) when
        // the rays run mostly along x and are offset along y.  Steps whose voxel lies
        // outside the image keep the reference's linear index into the image.
        const int nxy = g.nx * g.ny, npix = nxy * g.nz;
        const bool tr = imgT != nullptr && (t.flags & RAY_XMAJOR);
        const float* src = tr ? imgT : img;
        const int jstep = ray.jinc * g.nx, kstep = ray.kinc * nxy;
        const int tistep = tr ? ray.iinc * g.ny : ray.iinc, tjstep = tr ? ray.jinc : jstep;
        int li = ray.k * nxy + ray.j * g.nx + ray.i;
        int lt = ray.k * nxy + (tr ? ray.i * g.ny + ray.j : ray.j * g.nx + ray.i);
        for (int m0 = 0; m0 < ray.n; m0 += kUnroll) {
          float wv[kUnroll], pv[kUnroll];
          bool tv[kUnroll];
#pragma unroll
          for (int u = 0; u < kUnroll; u++) {
            RayIdx c;
            const int ltcur = lt;
            wv[u] = step_bf(ray, c, li, jstep, kstep);
            lt += (c.ci != ray.i) ? tistep : ((c.cj != ray.j) ? tjstep : kstep);
            const int lcur = c.li;
            bool ok;
            if (DIMS == 3) ok = (c.ci < g.nx) && (c.cj < g.ny) && (c.ck < g.nz) && lcur >= 0;
            else ok = (c.cj >= 0) && (c.ci >= 0) && lcur < nxy;
            ok = ok && (m0 + u < ray.n) && lcur < npix;
            const bool inside = c.ci >= 0 && c.cj >= 0 && c.ci < g.nx && c.cj < g.ny;
            tv[u] = ok;
            pv[u] = (inside ? src : img)[!ok ? 0 : (inside ? ltcur : lcur)];
          }
#pragma unroll
          for (int u = 0; u < kUnroll; u++) {
            const float term = (DIMS == 3) ? wv[u] * pv[u] : 1.0f * wv[u] * pv[u];
            const float nacc = acc + term;
            acc = tv[u] ? nacc : acc;
          }
        }
      } else if (PREC == APIRL_PREC_F64_SETUP && DIMS == 3) {
        // F64_SETUP: the setup in double from the float LOR (siddon_f64.h), rounded to
        // float; then the reference's float loop (Siddon.cpp:310-340: repeated
        // additions, the x / y / z tie order) with a Kahan-compensated float sum
        SetupXY64 xy;
        SetupZ64 zs;
        setup_xy_f64(g, t.p0x, t.p0y, t.vx, t.vy, xy);
        setup_z_f64(g, xy, za, t.vx, t.vy, (double)(zb - za), zs);
        const int n = abs(xy.i1 - xy.i0) + abs(xy.j1 - xy.j0) + abs(zs.k1 - zs.k0);
        float ax = (float)xy.ax, ay = (float)xy.ay, az = (float)zs.az, ac = (float)xy.amin;
        const float axu = (float)xy.axu, ayu = (float)xy.ayu, azu = (float)zs.azu;
        const float L = (float)zs.L;
        const long nxy = (long)g.nx * g.ny;
        int i = xy.i0, j = xy.j0, kk = zs.k0;
        float c = 0.f;
        for (int m = 0; m < n; m++) {
          const int ci = i, cj = j, ck = kk;
          float an;
          if (ax <= ay && ax <= az) {
            an = ax;
            i += xy.iinc;
            ax = ax + axu;
          } else if (ay <= ax && ay <= az) {
            an = ay;
            j += xy.jinc;
            ay = ay + ayu;
          } else {
            an = az;
            kk += zs.kinc;
            az = az + azu;
          }
          const float w = (an - ac) * L;
          ac = an;
          if (ci >= 0 && cj >= 0 && ck >= 0 && ci < g.nx && cj < g.ny && ck < g.nz) {
            const float y = w * img[(long)ck * nxy + (long)cj * g.nx + ci] - c;
            const float tt = acc + y;
            c = (tt - acc) - y;
            acc = tt;
          }
        }
      } else if (PREC == APIRL_PREC_KAHAN || PREC == APIRL_PREC_F64_SETUP) {
        float c = 0.f;
        for (int m = 0; m < ray.n; m++) {
          int ci, cj, ck;
          const float w = step(ray, ci, cj, ck);
          long nxy = (long)g.nx * g.ny;
          bool ok;
          long li;
          if (DIMS == 3) {
            ok = (ci < g.nx) && (cj < g.ny) && (ck < g.nz);
            li = (long)ck * nxy + (long)cj * g.nx + ci;
            ok = ok && li >= 0;
          } else {
            ok = (cj >= 0) && (ci >= 0);
            li = (long)cj * g.nx + ci;
            ok = ok && li < nxy;
          }
          if (ok) {
            const float y = w * img[li] - c;
            const float tt = acc + y;
            c = (tt - acc) - y;
            acc = tt;
          }
        }
      } else {
        // F64_ALPHA.  3D: the whole Siddon setup and traversal in double from the
        // float LOR (siddon_f64.h), summed in double in list order.  2D: the
        // reference's float setup (entry voxel, list length, the 2D clamps) with the
        // crossings, weights and sum in double.
        double dacc = 0.0;
        const long nxy = (long)g.nx * g.ny;
        if (DIMS == 3) {
          SetupXY64 xy;
          SetupZ64 zs;
          const double vz = (double)(zb - za);
          setup_xy_f64(g, t.p0x, t.p0y, t.vx, t.vy, xy);
          setup_z_f64(g, xy, za, t.vx, t.vy, vz, zs);
          const int n = abs(xy.i1 - xy.i0) + abs(xy.j1 - xy.j0) + abs(zs.k1 - zs.k0);
          double ax = xy.ax, ay = xy.ay, az = zs.az, ac = xy.amin;
          int i = xy.i0, j = xy.j0, kk = zs.k0;
          for (int m = 0; m < n; m++) {
            const int ci = i, cj = j, ck = kk;
            double w;
            if (ax <= ay && ax <= az) {
              w = (ax - ac) * zs.L;
              i += xy.iinc; ac = ax; ax += xy.axu;
            } else if (ay <= ax && ay <= az) {
              w = (ay - ac) * zs.L;
              j += xy.jinc; ac = ay; ay += xy.ayu;
            } else {
              w = (az - ac) * zs.L;
              kk += zs.kinc; ac = az; az += zs.azu;
            }
            if (ci >= 0 && cj >= 0 && ck >= 0 && ci < g.nx && cj < g.ny && ck < g.nz)
              dacc += w * (double)img[(long)ck * nxy + (long)cj * g.nx + ci];
          }
        } else {
          double ax = ray.ax, ay = ray.ay, ac = ray.ac;
          const double axu = ray.axu, ayu = ray.ayu, L = ray.L;
          int i = ray.i, j = ray.j;
          for (int m = 0; m < ray.n; m++) {
            const int ci = i, cj = j;
            const bool bx = ax <= ay;
            const double an = bx ? ax : ay;
            const double w = (an - ac) * L;
            ac = an;
            if (bx) { i += ray.iinc; ax += axu; }
            else { j += ray.jinc; ay += ayu; }
            if (cj >= 0 && ci >= 0 && ci < g.nx && cj < g.ny)
              dacc += w * (double)img[(long)cj * g.nx + ci];
          }
        }
        acc = (float)dacc;
      }
    }
    if (!FUSED) {
      if (valid) sino[idx] = acc;
    } else {
      // fused EM epilogue: est (+ additive), likelihood partial, ratio
      // (Sinogram3D.cpp:584-607, 830-870; Sinogram2D.cpp:234-266, 565-581)
      const long full = ((long)s * g.nang + a) * g.nr + rr;
      float est = valid ? acc : 0.f;
      if (fr.additive)
        est = est + fr.additive[fr.additive_local ? ((long)s * g.nang + k) * g.nr + rr : full];
      const float y = fr.meas[full];
      const bool lik_term = fr.lik_positive_only ? (est > 0.f) : (est != 0.f);
      if (lik_term) lik = (double)(y * logf(est) - est);
      fr.ratio[idx] = (y != 0.f && est != 0.f) ? y / est : 0.f;
    }
  }
  return lik;
}

template <int PREC, bool FUSED, int DIMS>
__global__ __launch_bounds__(kBlock) void project_kernel(
    Geo g, const RayXY* __restrict__ rays, const float2* __restrict__ zab,
    const float* __restrict__ img, float* __restrict__ sino, int nang_sub, int subset,
    int nsubsets, long nrays, FusedRatio fr, const int* __restrict__ zorder,
    const float* __restrict__ imgT) {
  // Rays are taken in order of their sinogram's mean axial coordinate (zorder) and
  // each XCD gets one contiguous share of that order: workgroups are dispatched to the
  // 8 XCDs round-robin, so block b runs on XCD b % 8, and the remap below gives XCD x
  // the x-th eighth of the z-sorted rays.  Its 4 MB L2 then serves one axial band of
  // the image instead of all of it.
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned xcd = b % kNumXcd, per = nb / kNumXcd, rem = nb % kNumXcd;
  const unsigned lb = xcd * per + (xcd < rem ? xcd : rem) + b / kNumXcd;
  const long lidx = (long)lb * kBlock + threadIdx.x;
  const double lik = project_ray<PREC, FUSED, DIMS>(g, rays, zab, img, sino, nang_sub, subset,
                                                    nsubsets, nrays, fr, zorder, imgT, lidx);
  if (FUSED && fr.lik_partials) {
    // deterministic block reduction of the likelihood partials (fixed tree)
    __shared__ double red[kBlock];
    red[threadIdx.x] = lik;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) fr.lik_partials[blockIdx.x] = red[0];
  }
}

// Ray-driven scatter backprojection with fp32 atomics.
//   BP_PLAIN : term = w * b; 3D skips b == 0 (SiddonProjector.cpp:496) with a full
//              bounds check (:541-548); 2D keeps every bin (:50-77) and writes
//              (w * b) * geom (:72).
//   BP_ONES  : b = 1 (sensitivity of a subset).
//   BP_MULT  : b = multiplicative sinogram bin (FULL layout).
//   BP_DIVIDE: 3D: skip meas == 0; term = (w*m)/e if e != 0 else w*m (:607-685) with the
//              upper-bound-only check of :666; 2D: skip e == 0 or m == 0 (:110);
//              term = (w*m)/e with the full check (:125-131).
// Hardware id of the XCD (accelerator complex die) running this wave: s_getreg of
// HW_REG_XCC_ID (hwreg 20, bits [3:0]) on gfx940+.
__device__ __forceinline__ int xcc_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (kNumXcd - 1);
}

template <int KIND, bool REPLICA>
__global__ __launch_bounds__(kBlock) void backproject_scatter_kernel(
    Geo g, const RayXY* __restrict__ rays, const float2* __restrict__ zab,
    const float* __restrict__ sino, const float* __restrict__ est, float* __restrict__ img,
    int nang_sub, int subset, int nsubsets, long nrays) {
  // Lanes take consecutive radial bins of one (sinogram, angle): the atomics of one
  // instruction then fall in few cache lines (measured faster than spreading the
  // lanes over sinograms, which puts every lane on its own line).
  const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  if (idx >= nrays) return;
  // With REPLICA, `img` holds 2 x 8 image copies: each wave adds into its own die's
  // copies, so the atomics of a cache line stay in one L2 instead of moving between
  // dies.  Rays running mostly along x (whose 64 lanes are offset along y) add their
  // in-image voxels into the die's transposed copy (y fastest), where one atomic
  // instruction of the wavefront touches a few cache lines instead of one line per
  // lane.  reduce_replicas_kernel folds the copies afterwards.
  float* imgT = img;
  if (REPLICA) {
    const long nv = (long)g.nx * g.ny * g.nz;
    imgT = img + (long)(kNumXcd + xcc_id()) * nv;
    img += (long)xcc_id() * nv;
  }
  int s, a, rr, k;
  ray_coords(g, idx, nang_sub, subset, nsubsets, s, a, rr, k);
  float b = 1.f, e = 0.f;
  if (KIND == BP_PLAIN) b = sino[idx];
  if (KIND == BP_MULT) b = sino[((long)s * g.nang + a) * g.nr + rr];
  if (KIND == BP_DIVIDE) {
    b = sino[idx];
    e = est[idx];
    if (b == 0.f) return;
    if (g.dims == 2 && e == 0.f) return;
  }
  if (g.dims == 3 && (KIND == BP_PLAIN || KIND == BP_MULT) && b == 0.f) return;
  const long nxy = (long)g.nx * g.ny;
  const long npix = nxy * g.nz;
  // every LOR of the bin (ns x nas when oversampled, SiddonProjector.cpp:491-548)
  for (int o = 0; o < g.ns; o++)
  for (int q = 0; q < g.nas; q++) {
  const RayXY t = rays[((long)a * g.nr + rr) * g.ns + o];
  float za = 0.f, zb = 0.f;
  if (g.dims == 3) {
    const float2 z = zab[(long)s * g.nas + q];
    za = z.x;
    zb = z.y;
  }
  Ray ray;
  const bool valid = g.dims == 3 ? setup_ray<3>(g, t, za, zb, ray) : setup_ray<2>(g, t, za, zb, ray);
  if (!valid) continue;
  const bool tr = REPLICA && (t.flags & RAY_XMAJOR);
  for (int m = 0; m < ray.n; m++) {
    int ci, cj, ck;
    const float w = step(ray, ci, cj, ck);
    bool ok;
    float term;
    if (KIND == BP_DIVIDE && g.dims == 3) {
      ok = (ck < g.nz) && (cj < g.ny) && (ci < g.nx);
      term = (e != 0.f) ? w * 1.0f * b / e : w * 1.0f * b;
    } else {
      ok = (ck < g.nz) && (cj < g.ny) && (ci < g.nx) && (ck >= 0) && (cj >= 0) && (ci >= 0);
      if (KIND == BP_DIVIDE) term = 1.0f * w * b / e;
      else if (g.dims == 2) term = w * b * 1.0f;
      else term = w * 1.0f * b;
    }
    if (ok) {
      const long li = (long)ck * nxy + (long)cj * g.nx + ci;
      // Only a voxel inside the image goes to the transposed copy.  DivideAndBackproject
      // tests the upper bounds only (SiddonProjector.cpp:666-667): there a negative ci
      // or cj names the voxel li of the image layout (as the reference does), which has
      // no transposed counterpart, and a negative ck gives li < 0, which is skipped.
      if (tr && ci >= 0 && cj >= 0 && ck >= 0)
        unsafeAtomicAdd(&imgT[(long)ck * nxy + (long)ci * g.ny + cj], term);
      else if (li >= 0 && li < npix) unsafeAtomicAdd(&img[li], term);
    }
  }
  }
}

__global__ __launch_bounds__(kBlock) void count_kernel(Geo g, const RayXY* __restrict__ rays,
                                                       const float2* __restrict__ zab,
                                                       int nang_sub, int subset, int nsubsets,
                                                       long nrays,
                                                       unsigned long long* __restrict__ out) {
  const long idx = (long)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long n = 0;
  if (idx < nrays) {
    int s, a, rr, k;
    ray_coords(g, idx, nang_sub, subset, nsubsets, s, a, rr, k);
    for (int o = 0; o < g.ns; o++)
      for (int q = 0; q < g.nas; q++) {
        const RayXY t = rays[((long)a * g.nr + rr) * g.ns + o];
        float za = 0.f, zb = 0.f;
        if (g.dims == 3) {
          za = zab[(long)s * g.nas + q].x;
          zb = zab[(long)s * g.nas + q].y;
        }
        Ray ray;
        const bool valid =
            g.dims == 3 ? setup_ray<3>(g, t, za, zb, ray) : setup_ray<2>(g, t, za, zb, ray);
        if (valid) n += (unsigned long long)ray.n;
      }
  }
  __shared__ unsigned long long red[kBlock];
  red[threadIdx.x] = n;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

__global__ __launch_bounds__(kBlock) void em_update_kernel(float* __restrict__ x,
                                                           const float* __restrict__ b,
                                                           const float* __restrict__ s,
                                                           long n, float thr, int strict) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float sv = s[i];
  const bool ok = strict ? (sv > thr) : (sv >= thr);
  x[i] = ok ? x[i] * b[i] / sv : 0.f;
}

// Per-bin value that DivideAndBackproject backprojects, for its gather form:
// 3D (SiddonProjector.cpp:607-685): m == 0 -> 0, e != 0 -> m / e, else m;
// 2D (:110-131): m == 0 or e == 0 -> 0, else m / e.
__global__ __launch_bounds__(kBlock) void divide_rule_kernel(const float* __restrict__ m,
                                                             const float* __restrict__ e,
                                                             float* __restrict__ out, long n,
                                                             int dims) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float mv = m[i], ev = e[i];
  float r;
  if (dims == 3) r = (mv == 0.f) ? 0.f : ((ev != 0.f) ? mv / ev : mv);
  else r = (mv == 0.f || ev == 0.f) ? 0.f : mv / ev;
  out[i] = r;
}

__global__ __launch_bounds__(kBlock) void force_positive_kernel(float* __restrict__ x, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n && x[i] < 0.f) x[i] = 0.f;
}

// Fixed-order sum of block partials (one block): deterministic for a given count.
__global__ __launch_bounds__(kBlock) void sum_partials_kernel(const double* __restrict__ p,
                                                              int n, double* __restrict__ out) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) acc += p[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Minimum of the non-zero values and maximum (Images.cpp:397-431).  Both are
// order-independent, so the result is exact.  out[0] = min_nz, out[1] = max; the
// reference's maximum starts from pixel 0, so a NaN there is kept (out[2] flags it).
__global__ __launch_bounds__(kBlock) void min_nz_max_kernel(const float* __restrict__ x, long n,
                                                            float* __restrict__ out) {
  __shared__ float smin[kBlock], smax[kBlock];
  float mn = FLT_MAX, mx = -FLT_MAX;
  for (long i = threadIdx.x; i < n; i += kBlock) {
    const float v = x[i];
    if (v != 0.f && v < mn) mn = v;
    if (v > mx) mx = v;
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + off]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smin[0];
    out[1] = (x[0] > smax[0] || x[0] != x[0]) ? x[0] : smax[0];
  }
}

__global__ __launch_bounds__(kBlock) void likelihood_kernel(const float* __restrict__ x,
                                                            const float* __restrict__ y, long n,
                                                            int positive_only,
                                                            double* __restrict__ partials) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const float e = x[i];
    if (positive_only ? (e > 0.f) : (e != 0.f)) v = (double)(y[i] * logf(e) - e);
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// img[v] += rep[0][v] + ... + rep[7][v] (fixed order), and clears the replicas for
// the next backprojection.
// img += the 8 image-layout copies rep[0..7] and the 8 transposed copies rep[8..15]
// (layout [z][x][y]), which are cleared for the next use.  One 32 x 32 tile of one
// slice per workgroup: the transposed sums pass through LDS, so every copy is read
// and cleared with coalesced accesses.  Summation order per voxel: the image-layout
// copies by die, then the transposed ones by die.
// The copies are read and cleared with device-scope accesses: every die's atomics and
// the next launch's see the same values whichever die's L2 last held a line.
__device__ __forceinline__ float rep_load(const float* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void rep_clear(float* p) {
  __hip_atomic_store(p, 0.f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBlock) void zero_floats_kernel(float* __restrict__ p, long n) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) rep_clear(p + i);
}

__global__ __launch_bounds__(kBlock) void reduce_replicas_kernel(float* __restrict__ img,
                                                                 float* __restrict__ rep, int nx,
                                                                 int ny) {
  __shared__ float tile[32][33];
  const long n = (long)nx * ny * gridDim.z;
  const long base = (long)blockIdx.z * nx * ny;
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int x = x0 + r, y = y0 + tx;
    float acc = 0.f;
    if (x < nx && y < ny) {
      const long i = base + (long)x * ny + y;
      float v[kNumXcd];
#pragma unroll
      for (int c = 0; c < kNumXcd; c++) v[c] = rep_load(&rep[(long)(kNumXcd + c) * n + i]);
#pragma unroll
      for (int c = 0; c < kNumXcd; c++) acc += v[c];
#pragma unroll
      for (int c = 0; c < kNumXcd; c++) rep_clear(&rep[(long)(kNumXcd + c) * n + i]);
    }
    tile[r][tx] = acc;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int y = y0 + r, x = x0 + tx;
    if (x < nx && y < ny) {
      const long i = base + (long)y * nx + x;
      float acc = 0.f, v[kNumXcd];
#pragma unroll
      for (int c = 0; c < kNumXcd; c++) v[c] = rep_load(&rep[(long)c * n + i]);
#pragma unroll
      for (int c = 0; c < kNumXcd; c++) acc += v[c];
#pragma unroll
      for (int c = 0; c < kNumXcd; c++) rep_clear(&rep[(long)c * n + i]);
      img[i] += acc + tile[tx][r];
    }
  }
}

// out[k][i][j] = in[k][j][i] through a 32 x 32 LDS tile (coalesced both ways).
__global__ __launch_bounds__(kBlock) void transpose_xy_kernel(const float* __restrict__ in,
                                                              float* __restrict__ out, int nx,
                                                              int ny) {
  __shared__ float tile[32][33];
  const int k = blockIdx.z;
  const long base = (long)k * nx * ny;
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int x = x0 + tx, y = y0 + r;
    if (x < nx && y < ny) tile[r][tx] = in[base + (long)y * nx + x];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int y = y0 + tx, x = x0 + r;
    if (x < nx && y < ny) out[base + (long)x * ny + y] = tile[tx][r];
  }
}

// ---- grouped (multi-job) launches --------------------------------------------
static_assert(kGroupBlock == kBlock, "grouped launches use the projector's block size");

// Grouped-launch helpers (also used by gather_kernels.hip through group_common.h).
#include "group_common.h"

__global__ __launch_bounds__(kBlock) void clear_group_kernel(BatchHeader* __restrict__ hdr,
                                                             const ClearJob* __restrict__ jobs) {
  const int n = hdr->count[PH_CLEAR];
  const long total = hdr->total[PH_CLEAR];
  const int grab = grab_size(hdr);
  for (;;) {
    const long v0 = grab_vb(&hdr->next[PH_CLEAR], grab);
    if (v0 >= total) break;
    for (long vb = v0; vb < v0 + grab && vb < total; vb++) {
      const ClearJob& J = jobs[find_job(jobs, n, vb)];
      const long i = (vb - J.first) * kBlock + threadIdx.x;
      if (i < J.n) J.p[i] = 0.f;
    }
  }
}

template <int DIMS>
__global__ __launch_bounds__(kBlock) void project_group_kernel(BatchHeader* __restrict__ hdr,
                                                               const ProjJob* __restrict__ jobs) {
  const int ph = DIMS == 2 ? PH_PROJ2 : PH_PROJ3;
  const int n = hdr->count[ph];
  const long total = hdr->total[ph];
  const int grab = grab_size(hdr);
  for (;;) {
    const long v0 = grab_vb(&hdr->next[ph], grab);
    if (v0 >= total) break;
    for (long vb = v0; vb < v0 + grab && vb < total; vb++) {
      const ProjJob& J = jobs[find_job(jobs, n, vb)];
      const long lidx = (vb - J.first) * kBlock + threadIdx.x;
      (void)project_ray<APIRL_PREC_REFERENCE, true, DIMS>(J.g, J.rays, J.zab, J.img, nullptr,
                                                          J.nang_sub, J.subset, J.nsubsets,
                                                          J.nrays, J.fr, J.zorder, nullptr, lidx);
    }
  }
}

__global__ __launch_bounds__(kBlock) void update_group_kernel(BatchHeader* __restrict__ hdr,
                                                              const UpdJob* __restrict__ jobs) {
  const int n = hdr->count[PH_UPDATE];
  const long total = hdr->total[PH_UPDATE];
  const int grab = grab_size(hdr);
  for (;;) {
    const long v0 = grab_vb(&hdr->next[PH_UPDATE], grab);
    if (v0 >= total) break;
    for (long vb = v0; vb < v0 + grab && vb < total; vb++) {
      const UpdJob& J = jobs[find_job(jobs, n, vb)];
      const long i = (vb - J.first) * kBlock + threadIdx.x;
      if (i < J.n) {
        const float sv = J.s[i];
        const bool ok = J.strict ? (sv > J.thr) : (sv >= J.thr);
        J.x[i] = ok ? J.x[i] * J.b[i] / sv : 0.f;
      }
    }
  }
}

inline unsigned grid_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }


// ---- Path forward projector (3D, APIRL_PREC_REFERENCE, one LOR per bin) ------------
//
// The rays of one (angle, radial bin) in every sinogram share their transaxial path:
// the reference's loop takes its x and y steps in the same order whatever its z steps
// (Siddon.cpp:318-339: a z step is taken only below both, and x before y on ties).
// project_paths_kernel takes one (angle, bin) per workgroup, first writes that path to
// LDS (the columns the loop visits and the alpha at which it leaves each, one thread
// running the loop's x / y additions), then takes the sinograms one per lane: each
// lane runs the reference's loop, taking a z step when its next z crossing is below
// the column's exit alpha and the path's next column otherwise, and adds w * voxel in
// the loop's order, so the bin is
// the reference's float sum, bit for bit (as project_kernel's).  Voxels come from a
// copy of the image with z fastest, where the lanes, at one column and neighbouring
// slices, read neighbouring words.

struct FwdStep {
  float ae;     // alpha at which the loop leaves this column by an x or y step
  int base;     // (j nx + i) nz: the column's first voxel in the z-fastest copy
  short i, j;   // the column (may lie outside the image at the path's ends)
  int run;      // columns inside the image from this one on, this one included
};
static_assert(sizeof(FwdStep) == 16, "FwdStep is 16 bytes");

// Path length kept per (angle, bin): the transaxial list plus a few columns (a rounding
// can trade a z step of the list for an x or y step).
__host__ __device__ inline int fwd_path_len(const Geo& g) { return g.nx + g.ny + 4; }
// LDS of project_paths_kernel: the list (one entry more) and the two crossing chains
inline size_t fwd_path_lds(const Geo& g) {
  return (size_t)(fwd_path_len(g) + 1) * (sizeof(FwdStep) + 2 * sizeof(float));
}

// out[(j nx + i) nz + k] = in[(k ny + j) nx + i] through a 32 (i) x 32 (k) LDS tile.
__global__ __launch_bounds__(kBlock) void transpose_zfast_kernel(const float* __restrict__ in,
                                                                 float* __restrict__ out, int nx,
                                                                 int ny, int nz) {
  __shared__ float tile[32][33];
  const int i0 = blockIdx.x * 32, j = blockIdx.y, k0 = blockIdx.z * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  const long nxy = (long)nx * ny;
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r, i = i0 + tx;
    if (k < nz && i < nx) tile[r][tx] = in[(long)k * nxy + (long)j * nx + i];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, k = k0 + tx;
    if (i < nx && k < nz) out[((long)j * nx + i) * nz + k] = tile[tx][r];
  }
}

template <bool FUSED>
__global__ __launch_bounds__(kBlock) void project_paths_kernel(
    Geo g, const RayXY* __restrict__ rays, const float2* __restrict__ zab,
    const float* __restrict__ img, const float* __restrict__ imgZ, float* __restrict__ sino,
    int nang_sub, int subset, int nsubsets, FusedRatio fr, int qmax) {
  extern __shared__ FwdStep sp[];
  __shared__ int s_qin[2];  // first and last list entries inside the image
  const int pair = blockIdx.x;
  const int ka = pair / g.nr, rr = pair % g.nr;
  const int a = subset + nsubsets * ka;
  const RayXY t = rays[(long)a * g.nr + rr];
  // The list has qmax + 1 entries (a fast batch that starts kPathUnroll columns before the
  // list's end reads the entry after it).  The loop's x crossings X[m] (X[0] = ax,
  // X[m + 1] = X[m] + axu, in float as the loop adds them) and y crossings Y[k] are
  // two chains, one lane each; the loop takes the x step at X[m] when X[m] <= the
  // pending Y, so it takes X[m] at step m + #{k: Y[k] < X[m]} and Y[k] at step
  // k + #{m: X[m] <= Y[k]} (a stable merge of the two sorted chains, x first on ties):
  // each step's entry is written by the crossing the loop takes there.
  float* xs = (float*)(sp + qmax + 1);
  float* ys = xs + (qmax + 1);
  if (threadIdx.x < 2) {
    const bool isx = threadIdx.x == 0;
    float v = isx ? t.ax : t.ay;
    const float u = isx ? t.axu : t.ayu;
    float* arr = isx ? xs : ys;
    for (int m = 0; m <= qmax; m++) {
      arr[m] = v;
      v = v + u;
    }
  }
  if (threadIdx.x == 0) {
    s_qin[0] = qmax + 1;
    s_qin[1] = -1;
  }
  __syncthreads();
  {
    const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
    for (int c = threadIdx.x; c < 2 * (qmax + 1); c += kBlock) {
      const bool isx = c <= qmax;
      const int m = isx ? c : c - (qmax + 1);
      const float v = isx ? xs[m] : ys[m];
      const float* other = isx ? ys : xs;
      int lo = 0, hi = qmax + 1;  // other[0 .. lo) come before v
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const float o = other[mid];
        const bool before = isx ? (o < v) : (o <= v);
        lo = before ? mid + 1 : lo;
        hi = before ? hi : mid;
      }
      const int q = m + lo;
      if (q <= qmax) {
        const int mx = isx ? m : lo, my = isx ? lo : m;
        const int i = t.i0 + mx * iinc, j = t.j0 + my * jinc;
        sp[q].ae = v;
        sp[q].i = (short)i;
        sp[q].j = (short)j;
        sp[q].base = (j * g.nx + i) * g.nz;
        // the entries inside the image are one run: i and j each move one way, so each
        // stays in range over one interval of entries
        if ((unsigned)i < (unsigned)g.nx && (unsigned)j < (unsigned)g.ny) {
          atomicMin(&s_qin[0], q);
          atomicMax(&s_qin[1], q);
        }
      }
    }
  }
  __syncthreads();
  {
    const int qa = s_qin[0], qb = s_qin[1];
    for (int q = threadIdx.x; q <= qmax; q += kBlock)
      sp[q].run = q >= qa && q <= qb ? qb - q + 1 : 0;
  }
  __syncthreads();
  const int nxy = g.nx * g.ny, npix = nxy * g.nz;
  double lik = 0.0;
  for (int s = threadIdx.x; s < g.nsino; s += kBlock) {
    const float2 zz = zab[s];
    Ray ray;
    const bool valid = setup_ray<3>(g, t, zz.x, zz.y, ray);
    float acc = 0.f;
    if (valid) {
      // Steps are taken kPathUnroll at a time: the batch's voxel loads are issued together,
      // then added in the loop's order.
      int q = 0, k = ray.k;
      FwdStep e = sp[0];
      float Zn = ray.az, ac = ray.ac;
      for (int m0 = 0; m0 < ray.n; m0 += kPathUnroll) {
        float wv[kPathUnroll], pv[kPathUnroll];
        // a batch whose kPathUnroll steps all stay in the image (the path's next kPathUnroll
        // columns inside; the slice inside and kPathUnroll - 1 z steps from the face it
        // moves towards) needs no read test; the wavefront takes it when every lane can
        const int kfar = k + (kPathUnroll - 1) * ray.kinc;
        const bool fast = (m0 + kPathUnroll <= ray.n) & (e.run >= kPathUnroll) & (k >= 0) &
                          (k < g.nz) & (kfar >= 0) & (kfar < g.nz);
        if (__all(fast)) {
#pragma unroll
          for (int u = 0; u < kPathUnroll; u++) {
            const bool zst = Zn < e.ae;
            const float an = zst ? Zn : e.ae;
            wv[u] = (an - ac) * ray.L;
            ac = an;
            pv[u] = imgZ[(unsigned)(e.base + k)];  // inside the image: non-negative
            Zn = zst ? Zn + ray.azu : Zn;
            k += zst ? ray.kinc : 0;
            q += zst ? 0 : 1;
            e = sp[q];
          }
#pragma unroll
          for (int u = 0; u < kPathUnroll; u++) acc = acc + wv[u] * pv[u];
          continue;
        }
        bool tv[kPathUnroll];
#pragma unroll
        for (int u = 0; u < kPathUnroll; u++) {
          const bool zst = Zn < e.ae;
          const float an = zst ? Zn : e.ae;
          wv[u] = (an - ac) * ray.L;
          ac = an;
          // the reference's read test (upper bounds and a non-negative linear index)
          const int ci = e.i, cj = e.j;
          const int li = k * nxy + cj * g.nx + ci;
          const bool ok = (ci < g.nx) & (cj < g.ny) & (k < g.nz) & (li >= 0) & (li < npix) &
                          (m0 + u < ray.n);
          const bool inside = (ci >= 0) & (cj >= 0);
          tv[u] = ok;
          pv[u] = ok ? (inside ? imgZ[e.base + k] : img[li]) : 0.f;
          Zn = zst ? Zn + ray.azu : Zn;
          k += zst ? ray.kinc : 0;
          q += zst ? 0 : 1;
          e = sp[q < qmax ? q : qmax - 1];
        }
#pragma unroll
        for (int u = 0; u < kPathUnroll; u++) {
          const float term = wv[u] * pv[u];
          const float nacc = acc + term;
          acc = tv[u] ? nacc : acc;
        }
      }
    }
    const long idx = ((long)s * nang_sub + ka) * g.nr + rr;
    if (!FUSED) {
      if (valid) sino[idx] = acc;
    } else {
      // fused EM epilogue, as project_ray's (Sinogram3D.cpp:584-607, 830-870)
      const long full = ((long)s * g.nang + a) * g.nr + rr;
      float est = valid ? acc : 0.f;
      if (fr.additive)
        est = est + fr.additive[fr.additive_local ? ((long)s * g.nang + ka) * g.nr + rr : full];
      const float y = fr.meas[full];
      const bool lik_term = fr.lik_positive_only ? (est > 0.f) : (est != 0.f);
      if (lik_term) lik += (double)(y * logf(est) - est);
      fr.ratio[idx] = (y != 0.f && est != 0.f) ? y / est : 0.f;
    }
  }
  if (FUSED && fr.lik_partials) {
    // deterministic block reduction of the likelihood partials (fixed tree)
    __shared__ double red[kBlock];
    red[threadIdx.x] = lik;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) fr.lik_partials[blockIdx.x] = red[0];
  }
}

}  // namespace

bool path_forward_ok(const apirl_plan* p, int precision) {
  const Geo& g = p->geo;
  return g.dims == 3 && precision == APIRL_PREC_REFERENCE && g.ns * g.nas == 1 && g.walk24 &&
         p->d_imgT != nullptr && fwd_path_lds(g) <= 48 * 1024;
}

int project_partials(const apirl_plan* p, int subset, int nsubsets, bool path) {
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  if (path) return nang_sub * g.nr;
  return (int)grid_for((long)g.nsino * nang_sub * g.nr);
}

hipError_t launch_project(const apirl_plan* p, const float* img, float* sino, int subset,
                          int nsubsets, int precision, const FusedRatio* fused, int* nblocks_out,
                          hipStream_t st) {
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  const unsigned grid = grid_for(nrays);
  if (nblocks_out) *nblocks_out = (int)grid;
  if (nrays == 0) return hipSuccess;
  FusedRatio fr = fused ? *fused : FusedRatio{};
  const bool f = fused != nullptr;
  if (path_forward_ok(p, precision)) {
    // the path forward projector: the image with z fastest (into the plan's
    // transposed-image buffer), one workgroup per (angle, bin) of the subset
    const int npair = nang_sub * g.nr, qmax = fwd_path_len(g);
    if (nblocks_out) *nblocks_out = npair;
    const dim3 tg((g.nx + 31) / 32, g.ny, (g.nz + 31) / 32);
    hipLaunchKernelGGL(transpose_zfast_kernel, tg, dim3(kBlock), 0, st, img, p->d_imgT, g.nx,
                       g.ny, g.nz);
    const size_t lds = fwd_path_lds(g);  // 12.4 KB for 256 x 256
    if (f)
      hipLaunchKernelGGL(project_paths_kernel<true>, dim3(npair), dim3(kBlock), lds, st, g,
                         p->d_rays, p->d_zab, img, p->d_imgT, sino, nang_sub, subset, nsubsets, fr,
                         qmax);
    else
      hipLaunchKernelGGL(project_paths_kernel<false>, dim3(npair), dim3(kBlock), lds, st, g,
                         p->d_rays, p->d_zab, img, p->d_imgT, sino, nang_sub, subset, nsubsets, fr,
                         qmax);
    return hipGetLastError();
  }
  const float* imgT = nullptr;
  if (p->d_imgT && precision == APIRL_PREC_REFERENCE) {
    const dim3 tg((g.nx + 31) / 32, (g.ny + 31) / 32, g.nz);
    hipLaunchKernelGGL(transpose_xy_kernel, tg, dim3(kBlock), 0, st, img, p->d_imgT, g.nx, g.ny);
    imgT = p->d_imgT;
  }
#define APIRL_LAUNCH_PROJ3(PR, D)                                                           \
  if (f)                                                                                   \
    hipLaunchKernelGGL((project_kernel<PR, true, D>), dim3(grid), dim3(kBlock), 0, st, g,  \
                       p->d_rays, p->d_zab, img, sino, nang_sub, subset, nsubsets, nrays, fr, \
                       p->d_zorder, imgT);                                                 \
  else                                                                                     \
    hipLaunchKernelGGL((project_kernel<PR, false, D>), dim3(grid), dim3(kBlock), 0, st, g, \
                       p->d_rays, p->d_zab, img, sino, nang_sub, subset, nsubsets, nrays, fr, \
                       p->d_zorder, imgT);
#define APIRL_LAUNCH_PROJ(PR)            \
  if (g.dims == 3) {                     \
    APIRL_LAUNCH_PROJ3(PR, 3)            \
  } else {                               \
    APIRL_LAUNCH_PROJ3(PR, 2)            \
  }
  switch (precision) {
    case APIRL_PREC_KAHAN: APIRL_LAUNCH_PROJ(APIRL_PREC_KAHAN); break;
    case APIRL_PREC_F64_ALPHA: APIRL_LAUNCH_PROJ(APIRL_PREC_F64_ALPHA); break;
    case APIRL_PREC_F64_SETUP: APIRL_LAUNCH_PROJ(APIRL_PREC_F64_SETUP); break;
    default: APIRL_LAUNCH_PROJ(APIRL_PREC_REFERENCE); break;
  }
#undef APIRL_LAUNCH_PROJ3
#undef APIRL_LAUNCH_PROJ
  return hipGetLastError();
}

hipError_t launch_backproject(const apirl_plan* p, const float* sino, const float* est,
                              float* img, int subset, int nsubsets, int kind, int mode,
                              hipStream_t st) {
  if (mode != APIRL_BP_SCATTER) return hipErrorNotSupported;
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  if (nrays == 0) return hipSuccess;
  const unsigned grid = grid_for(nrays);
  float* rep = p->d_replicas;
  float* target = rep ? rep : img;
#define APIRL_LAUNCH_BP(K)                                                                   \
  if (rep)                                                                                   \
    hipLaunchKernelGGL((backproject_scatter_kernel<K, true>), dim3(grid), dim3(kBlock), 0, st, \
                       g, p->d_rays, p->d_zab, sino, est, target, nang_sub, subset, nsubsets,  \
                       nrays);                                                              \
  else                                                                                       \
    hipLaunchKernelGGL((backproject_scatter_kernel<K, false>), dim3(grid), dim3(kBlock), 0,   \
                       st, g, p->d_rays, p->d_zab, sino, est, target, nang_sub, subset,       \
                       nsubsets, nrays);
  switch (kind) {
    case BP_ONES: APIRL_LAUNCH_BP(BP_ONES); break;
    case BP_MULT: APIRL_LAUNCH_BP(BP_MULT); break;
    case BP_DIVIDE: APIRL_LAUNCH_BP(BP_DIVIDE); break;
    default: APIRL_LAUNCH_BP(BP_PLAIN); break;
  }
#undef APIRL_LAUNCH_BP
  if (rep)
    hipLaunchKernelGGL(reduce_replicas_kernel, dim3((g.nx + 31) / 32, (g.ny + 31) / 32, g.nz),
                       dim3(kBlock), 0, st, img, rep, g.nx, g.ny);
  return hipGetLastError();
}

hipError_t launch_count_intersections(const apirl_plan* p, int subset, int nsubsets,
                                      unsigned long long* d_out, hipStream_t st) {
  const Geo& g = p->geo;
  const int nang_sub = subset_num_angles(g.nang, nsubsets, subset);
  const long nrays = (long)g.nsino * nang_sub * g.nr;
  if (nrays == 0) return hipSuccess;
  hipLaunchKernelGGL(count_kernel, dim3(grid_for(nrays)), dim3(kBlock), 0, st, g, p->d_rays,
                     p->d_zab, nang_sub, subset, nsubsets, nrays, d_out);
  return hipGetLastError();
}

hipError_t launch_zero_replicas(float* rep, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_floats_kernel, dim3(grid_for((long)n)), dim3(kBlock), 0, st, rep,
                     (long)n);
  return hipGetLastError();
}

hipError_t launch_em_update(float* x, const float* b, const float* s, size_t n, float thr,
                            int strict, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(em_update_kernel, dim3(grid_for((long)n)), dim3(kBlock), 0, st, x, b, s,
                     (long)n, thr, strict);
  return hipGetLastError();
}

hipError_t launch_divide_rule(const float* meas, const float* est, float* out, size_t n,
                              int dims, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(divide_rule_kernel, dim3(grid_for((long)n)), dim3(kBlock), 0, st, meas,
                     est, out, (long)n, dims);
  return hipGetLastError();
}

hipError_t launch_force_positive(float* x, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(force_positive_kernel, dim3(grid_for((long)n)), dim3(kBlock), 0, st, x,
                     (long)n);
  return hipGetLastError();
}

// The likelihood's block partials of apirl_project_ratio_noalloc: 2^19 doubles
// (4 MiB) per GPU, resident until the process ends.  Every block writes its own entry
// before sum_partials_kernel reads them, so nothing needs clearing; calls that overlap
// on one GPU share the array and race.
static __device__ __align__(256) double scratch[kNoallocPartials];

hipError_t noalloc_scratch_address(double** out) {
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  *out = e == hipSuccess ? static_cast<double*>(ptr) : nullptr;
  return e;
}

hipError_t launch_sum_partials(const double* partials, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kBlock), 0, st, partials, n, out);
  return hipGetLastError();
}

hipError_t launch_likelihood(const float* x, const float* y, long n, int positive_only,
                             double* partials, double* out, hipStream_t st) {
  if (n <= 0) {
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kBlock), 0, st, partials, 0, out);
    return hipGetLastError();
  }
  const unsigned grid = grid_for(n);
  hipLaunchKernelGGL(likelihood_kernel, dim3(grid), dim3(kBlock), 0, st, x, y, n,
                     positive_only, partials);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kBlock), 0, st, partials, (int)grid,
                     out);
  return hipGetLastError();
}

hipError_t launch_min_nz_max(const float* x, size_t n, float* d_out2, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(min_nz_max_kernel, dim3(1), dim3(kBlock), 0, st, x, (long)n, d_out2);
  return hipGetLastError();
}

hipError_t launch_group_clear(BatchHeader* hdr, const ClearJob* jobs, int grid,
                              hipStream_t st) {
  hipLaunchKernelGGL(clear_group_kernel, dim3(grid), dim3(kBlock), 0, st, hdr, jobs);
  return hipGetLastError();
}

hipError_t launch_group_project(BatchHeader* hdr, const ProjJob* jobs2,
                                const ProjJob* jobs3, int grid, hipStream_t st) {
  hipLaunchKernelGGL(project_group_kernel<2>, dim3(grid), dim3(kBlock), 0, st, hdr, jobs2);
  hipLaunchKernelGGL(project_group_kernel<3>, dim3(grid), dim3(kBlock), 0, st, hdr, jobs3);
  return hipGetLastError();
}

hipError_t launch_group_update(BatchHeader* hdr, const UpdJob* jobs, int grid,
                               hipStream_t st) {
  hipLaunchKernelGGL(update_group_kernel, dim3(grid), dim3(kBlock), 0, st, hdr, jobs);
  return hipGetLastError();
}

}  // namespace apirl
