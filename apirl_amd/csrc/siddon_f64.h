// Siddon setup of a 3D LOR carried out in double: the statements of
// recon/src/Siddon.cpp:13-286 (alpha range of the square FOV with its [0,1] clamps,
// entry / exit voxels by floor(), list length |di|+|dj|+|dk|, first crossings and
// increments) evaluated in double from the float LOR (P1 = p0, P2 = p0 + v, both as
// the reference computes them in float).  Used by the APIRL_PREC_F64 and
// APIRL_PREC_F64_SETUP modes of the projector and of the gather backprojector, and on
// the host for the F64_SETUP tables of the gather's column kernel.
#ifndef APIRL_SIDDON_F64_H
#define APIRL_SIDDON_F64_H

#include <cfloat>
#include <cmath>

#include "apirl_internal.h"

namespace apirl {

struct SetupXY64 {
  double amin, amax, ax, ay, axu, ayu;
  int i0, j0, i1, j1, iinc, jinc;
};

struct SetupZ64 {
  double az, azu, L;
  int k0, k1, kinc;
};

__host__ __device__ inline void setup_xy_f64(const Geo& g, double p0x, double p0y, double vx,
                                             double vy, SetupXY64& s) {
  const double dx = g.dx, dy = g.dy;
  const double R = (dx * g.nx) / 2;
  double amin, amax;
  if (vx == 0) {
    const double a1 = (-R - p0y) / vy, a2 = (R - p0y) / vy;
    amin = a1 < a2 ? a1 : a2;
    amax = a1 < a2 ? a2 : a1;
  } else if (vy == 0) {
    const double a1 = (-R - p0x) / vx, a2 = (R - p0x) / vx;
    amin = a1 < a2 ? a1 : a2;
    amax = a1 < a2 ? a2 : a1;
  } else {
    const double ax1 = (-R - p0x) / vx, ay1 = (-R - p0y) / vy;
    const double ax2 = (R - p0x) / vx, ay2 = (R - p0y) / vy;
    const double lo1 = ax1 < ax2 ? ax1 : ax2, lo2 = ay1 < ay2 ? ay1 : ay2;
    const double hi1 = ax1 > ax2 ? ax1 : ax2, hi2 = ay1 > ay2 ? ay1 : ay2;
    amin = lo1 > lo2 ? lo1 : lo2;
    amax = hi1 < hi2 ? hi1 : hi2;
  }
  if (amin < 0 || amin > 1) amin = 0;
  if (amax > 1 || amax < 0) amax = 1;
  s.amin = amin;
  s.amax = amax;
  s.i0 = (int)floor((p0x + vx * amin + R) / dx);
  s.j0 = (int)floor((p0y + vy * amin + R) / dy);
  s.i1 = (int)floor((p0x + vx * amax + R) / dx);
  s.j1 = (int)floor((p0y + vy * amax + R) / dy);
  s.iinc = vx > 0 ? 1 : (vx < 0 ? -1 : 0);
  s.jinc = vy > 0 ? 1 : (vy < 0 ? -1 : 0);
  s.axu = fabs(dx / vx);
  s.ayu = fabs(dy / vy);
  if (vx > 0) s.ax = (-R + (s.i0 + s.iinc) * dx - p0x) / vx;
  else if (vx < 0) s.ax = (-R + s.i0 * dx - p0x) / vx;
  else s.ax = DBL_MAX;
  if (vy > 0) s.ay = (-R + (s.j0 + s.jinc) * dy - p0y) / vy;
  else if (vy < 0) s.ay = (-R + s.j0 * dy - p0y) / vy;
  else s.ay = DBL_MAX;
}

__host__ __device__ inline void setup_z_f64(const Geo& g, const SetupXY64& xy, double p0z,
                                            double vx, double vy, double vz, SetupZ64& s) {
  const double dz = g.dz;
  const double offz = dz / 2;
  s.k0 = (int)floor((p0z + vz * xy.amin - offz) / dz);
  s.k1 = (int)floor((p0z + vz * xy.amax - offz) / dz);
  s.kinc = vz > 0 ? 1 : (vz < 0 ? -1 : 0);
  s.azu = fabs(dz / vz);
  if (vz > 0) s.az = (offz + (s.k0 + s.kinc) * dz - p0z) / vz;
  else if (vz < 0) s.az = (offz + s.k0 * dz - p0z) / vz;
  else s.az = DBL_MAX;
  s.L = sqrt(vx * vx + vy * vy + vz * vz);
}

}  // namespace apirl

#endif
