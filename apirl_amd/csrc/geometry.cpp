// Host-side geometry: sinogram grids, LOR endpoints, ring tables and the
// transaxial part of the Siddon setup, evaluated in float with the reference's
// operation order (this file is compiled with -ffp-contract=off; the x86-64
// baseline has no FMA either, so every product and sum is rounded as written).
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstdlib>

#include "apirl_internal.h"
#include "siddon_f64.h"

namespace apirl {

namespace {
constexpr float kDegToRad = 0.01745329251994329576f;  // utils/inc/Geometry.h:12

inline float min_ref(float a, float b) { return (b < a) ? b : a; }  // std::min
inline float max_ref(float a, float b) { return (a < b) ? b : a; }  // std::max
}  // namespace

int subset_num_angles(int nang, int nsubsets, int subset) {
  // Sinogram3DCylindricalPet.cpp:168-171
  int n = (int)floorf((float)nang / (float)nsubsets);
  if ((nang % nsubsets) > subset) n++;
  return n;
}

// LOR endpoints of bin (ang, r): Sinogram2DinCylindrical3Dpet.cpp:249-266 (the 3D
// form :216-229 is the same in x/y and never rejects).  The reference compiles to
// glibc sincosf/sqrtf on float arguments; the same calls are made here, on the host.
static bool lor_points(float ang_deg, float r, float rscanner, float rfov, bool check_fov,
                       float* p1x, float* p1y, float* p2x, float* p2y) {
  float rad = ang_deg * kDegToRad;
  float aux = sqrtf(rscanner * rscanner - r * r);
  if (check_fov && r > rfov) return false;
  float s, c;
  sincosf(rad, &s, &c);
  *p1x = r * c + s * aux;
  *p1y = r * s - c * aux;
  *p2x = r * c - s * aux;
  *p2y = r * s + c * aux;
  return true;
}

void build_ray_table(const apirl_image_desc& img, const apirl_sino_desc& sino,
                     const std::vector<float>& ang, const std::vector<float>& r,
                     std::vector<RayXY>& out) {
  const bool is2d = sino.dims == 2;
  std::vector<float4> pts((size_t)sino.nang * sino.nr);
  std::vector<char> ok(pts.size());
  for (int a = 0; a < sino.nang; a++)
    for (int j = 0; j < sino.nr; j++) {
      float4& q = pts[(size_t)a * sino.nr + j];
      q = float4{0, 0, 0, 0};
      ok[(size_t)a * sino.nr + j] = lor_points(ang[a], r[j], sino.rscanner_mm, sino.rfov_mm, is2d,
                                               &q.x, &q.y, &q.z, &q.w);
    }
  build_ray_table_pts(img, is2d, pts, ok, out);
}

namespace {
constexpr float kMmrRs = 328.0f, kMmrDoi = 9.6f;

// Sinogram2DinSiemensMmr.cpp:27-50: bins of one detector element on the cylinder,
// r = (Rs + DOI cos(lr / Rs)) sin(lr / Rs) with lr the arc length of the bin centre;
// angles a * 180 / nang, odd bins half an angular step less.
void mmr_grid(int nang, int nr, std::vector<float>& r, std::vector<float>& ang, float& step) {
  const float bin0 = 4.0891f / 2.0f;
  const float bin = nr == 344 ? bin0 : bin0 * 344 / nr;
  r.resize(nr);
  ang.resize(nang);
  for (int j = 0; j < nr; j++) {
    const float lr = bin / 2 + (bin * (j - (float)(nr / 2)));
    r[j] = (kMmrRs + kMmrDoi * cosf(lr / kMmrRs)) * sinf(lr / kMmrRs);
  }
  const float inc = (float)180.0f / nang;
  for (int a = 0; a < nang; a++) ang[a] = a * inc;
  step = nang > 1 ? ang[1] - ang[0] : 0.f;
}
}  // namespace

void mmr_radial_grid(int nang, int nr, std::vector<float>& r, std::vector<float2>& angcs_odd) {
  std::vector<float> ang;
  float step;
  mmr_grid(nang, nr, r, ang, step);
  angcs_odd.resize(nang);
  for (int a = 0; a < nang; a++) {
    const double rad = ((double)ang[a] - step / 2.0) * (double)kDegToRad;
    angcs_odd[a].x = (float)cos(rad);
    angcs_odd[a].y = (float)sin(rad);
  }
}

void build_radial_lut(const std::vector<float>& r, std::vector<int2>& lut, float& rl0,
                      float& inv_h) {
  const int nr = (int)r.size();
  double dmin = 1e30;
  for (int j = 1; j < nr; j++) dmin = std::min(dmin, (double)r[j] - r[j - 1]);
  const double h = nr > 1 ? std::max(dmin / 2, 1e-3) : 1.0;
  const double lo = (double)r[0] - 2 * h, hi = (double)r[nr - 1] + 2 * h;
  const int n = (int)std::ceil((hi - lo) / h) + 1;
  lut.resize(n);
  int first = 0, last = -1;
  for (int k = 0; k < n; k++) {
    const double x = lo + k * h;
    while (first < nr && r[first] < x) first++;
    while (last + 1 < nr && r[last + 1] <= x) last++;
    lut[k].x = first;
    lut[k].y = last;
  }
  rl0 = (float)lo;
  inv_h = (float)(1.0 / h);
}

void mmr_lor_points(int nang, int nr, std::vector<float4>& pts) {
  // Sinogram2DinSiemensMmr.cpp:175-192 (getPointsFromLor): the LOR of bin (a, j) at
  // distance r[j] from the centre ends on the radius Rs + DOI.
  std::vector<float> r, ang;
  float step;
  mmr_grid(nang, nr, r, ang, step);
  const float Re = kMmrRs + kMmrDoi;
  pts.resize((size_t)nang * nr);
  for (int a = 0; a < nang; a++)
    for (int j = 0; j < nr; j++) {
      const float rad = (ang[a] - (j % 2) * step / 2) * kDegToRad;
      const float aux = sqrtf(Re * Re - r[j] * r[j]);
      const float c = cosf(rad), sn = sinf(rad);
      float4& q = pts[(size_t)a * nr + j];
      q.x = r[j] * c + sn * aux;
      q.y = r[j] * sn - c * aux;
      q.z = r[j] * c - sn * aux;
      q.w = r[j] * sn + c * aux;
    }
}

void mmr_subset_lor_points(int nang, int nr, std::vector<float4>& pts) {
  // Sinogram3DSiemensMmr::getSubset (Sinogram3DSiemensMmr.cpp:150-195) holds
  // Sinogram2DinCylindrical3Dpet sinograms on the radius 328 mm (its segments are
  // made inside the cylindrical base constructor) with the angles and the
  // arc-corrected radial positions of the full mMR sinogram copied in; their LORs are
  // the cylindrical 3D form (Sinogram2DinCylindrical3Dpet.cpp:216-229).
  std::vector<float> r, ang;
  float step;
  mmr_grid(nang, nr, r, ang, step);
  pts.resize((size_t)nang * nr);
  for (int a = 0; a < nang; a++)
    for (int j = 0; j < nr; j++) {
      float4& q = pts[(size_t)a * nr + j];
      lor_points(ang[a], r[j], kMmrRs, 0.f, false, &q.x, &q.y, &q.z, &q.w);
    }
}

void build_ray_table_pts(const apirl_image_desc& img, bool is2d, const std::vector<float4>& pts,
                         const std::vector<char>& ok, std::vector<RayXY>& out) {
  const float dx = img.dx, dy = img.dy;
  const float R = (dx * img.nx) / 2;
  out.resize(pts.size());
  for (size_t n = 0; n < pts.size(); n++) {
    {
      RayXY& t = out[n];
      t = RayXY{};
      const float p1x = pts[n].x, p1y = pts[n].y, p2x = pts[n].z, p2y = pts[n].w;
      if (!ok[n]) {
        t.flags = ray_flags(0, 0, 0, 0);
        continue;
      }
      const float p0x = p1x, p0y = p1y, vx = p2x - p1x, vy = p2y - p1y;
      // Siddon.cpp:83-145 / :430-491
      float amin, amax;
      if (vx == 0) {
        float a1 = (-R - p0y) / vy, a2 = (R - p0y) / vy;
        if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
      } else if (vy == 0) {
        float a1 = (-R - p0x) / vx, a2 = (R - p0x) / vx;
        if (a1 < a2) { amin = a1; amax = a2; } else { amin = a2; amax = a1; }
      } else {
        float ax1 = (-R - p0x) / vx, ay1 = (-R - p0y) / vy;
        float ax2 = (R - p0x) / vx, ay2 = (R - p0y) / vy;
        amin = max_ref(min_ref(ax1, ax2), min_ref(ay1, ay2));
        amax = min_ref(max_ref(ax1, ax2), max_ref(ay1, ay2));
      }
      if (amin < 0 || amin > 1) amin = 0;
      if (amax > 1 || amax < 0) amax = 1;
      float x1 = p0x + vx * amin, y1 = p0y + vy * amin;
      float x2 = p0x + vx * amax, y2 = p0y + vy * amax;
      int imin = (int)floorf((x1 + R) / dx), jmin = (int)floorf((y1 + R) / dy);
      int imax = (int)floorf((x2 + R) / dx), jmax = (int)floorf((y2 + R) / dy);
      bool rejected = false;
      if (is2d) {  // Siddon.cpp:515-519 (the 3D form has no such test)
        rejected = (imin < 0 && imax < 0) || (jmin < 0 && jmax < 0) ||
                   (imin >= img.nx && imax >= img.nx) || (jmin >= img.ny && jmax >= img.ny);
      }
      int iinc = vx > 0 ? 1 : (vx < 0 ? -1 : 0);
      int jinc = vy > 0 ? 1 : (vy < 0 ? -1 : 0);
      float ax, ay;
      if (vx > 0) ax = (-R + (imin + iinc) * dx - p0x) / vx;
      else if (vx < 0) ax = (-R + imin * dx - p0x) / vx;
      else ax = FLT_MAX;
      if (vy > 0) ay = (-R + (jmin + jinc) * dy - p0y) / vy;
      else if (vy < 0) ay = (-R + jmin * dy - p0y) / vy;
      else ay = FLT_MAX;
      if (is2d) {  // the 2D form clamps a negative first crossing (:596-606); 3D does not
        if (ax < 0) ax = amin;
        if (ay < 0) ay = amin;
      }
      float ex = (p0x + vx) - p0x, ey = (p0y + vy) - p0y;
      t.p0x = p0x; t.p0y = p0y; t.vx = vx; t.vy = vy;
      t.amin = amin; t.amax = amax;
      t.ax = ax; t.ay = ay;
      t.axu = fabsf(dx / vx);
      t.ayu = fabsf(dy / vy);
      t.lxy = ex * ex + ey * ey;
      t.i0 = imin; t.j0 = jmin;
      t.nxy = rejected ? 0 : std::abs(imax - imin) + std::abs(jmax - jmin);
      t.flags = ray_flags(iinc, jinc, 1, rejected) | (fabsf(vx) > fabsf(vy) ? RAY_XMAJOR : 0);
    }
  }
}

void build_z_table(const apirl_sino_desc& sino, std::vector<float2>& zab) {
  // SiddonProjector.cpp:335-336,394-395 with one axial sample per bin:
  // P.Z = zmean - w/2 + dz/2 + 0*dz, dz = w / 1.
  zab.resize(sino.nsino);
  const float w = sino.ring_width_mm;
  const float dzs = w / 1;
  for (int s = 0; s < sino.nsino; s++) {
    float z1 = sino.z1_mean_mm[s], z2 = sino.z2_mean_mm[s];
    zab[s].x = z1 - w / 2 + dzs / 2 + 0 * dzs;
    zab[s].y = z2 - w / 2 + dzs / 2 + 0 * dzs;
  }
}

void build_subsample_r(const std::vector<float>& r, int ns, std::vector<float>& r_sub) {
  // r of sub-LOR o of bin j: r - inc/2 + inc/ns * o + inc/(2 ns), inc the bin's width
  // from its neighbours (Sinogram2D.cpp:172-187), in float as written
  const int nr = (int)r.size();
  r_sub.resize((size_t)nr * ns);
  for (int j = 0; j < nr; j++) {
    float inc;
    if (j > 0 && j < nr - 1) inc = (r[j + 1] - r[j - 1]) / 2;
    else if (j == 0) inc = nr > 1 ? r[1] - r[0] : 0.f;
    else inc = r[nr - 1] - r[nr - 2];
    for (int o = 0; o < ns; o++)
      r_sub[(size_t)j * ns + o] = r[j] - (inc / 2) + inc / ns * o + (inc / (2 * ns));
  }
}

void build_z_table_os(const std::vector<float>& z1m, const std::vector<float>& z2m,
                      float ring_width, int nas, std::vector<float2>& zab) {
  // P.Z = zmean - w/2 + dz/2 + p dz, dz = w / nas (SiddonProjector.cpp:335-336, 394-395)
  const size_t ns = z1m.size();
  zab.resize(ns * nas);
  const float w = ring_width;
  const float dzs = w / nas;
  for (size_t s = 0; s < ns; s++)
    for (int q = 0; q < nas; q++) {
      zab[s * nas + q].x = z1m[s] - w / 2 + dzs / 2 + q * dzs;
      zab[s * nas + q].y = z2m[s] - w / 2 + dzs / 2 + q * dzs;
    }
}

void build_gather_tables(const apirl_plan& p, std::vector<float2>& angcs,
                         std::vector<GatherSeg>& segs, float& r0, float& dr, float& rmargin,
                         bool& ok) {
  const int nang = p.geo.nang, nr = p.geo.nr;
  angcs.resize(nang);
  for (int a = 0; a < nang; a++) {
    const double rad = (double)p.ang[a] * (double)kDegToRad;
    angcs[a].x = (float)cos(rad);
    angcs[a].y = (float)sin(rad);
  }
  r0 = p.r[0];
  dr = nr > 1 ? (p.r[nr - 1] - p.r[0]) / (float)(nr - 1) : 1.f;
  double dev = 0;
  for (int j = 0; j < nr; j++) dev = std::max(dev, std::fabs((double)p.r[j] - (r0 + (double)j * dr)));
  // candidate margin: float error of the LOR endpoints (~1e-4 mm) plus grid deviation
  rmargin = (float)(0.01 + dev);
  ok = nr >= 1 && dr > 0 && dev < 0.25 * dr;
  segs.clear();
  if (p.geo.dims != 3) return;
  const int ns = p.geo.nsino;
  int first = 0;
  for (int s = 1; s <= ns; s++) {
    const bool split = (s == ns) ||
                       (0.5f * (p.zab[s].x + p.zab[s].y) <= 0.5f * (p.zab[s - 1].x + p.zab[s - 1].y));
    if (!split) continue;
    GatherSeg g{};
    g.first = first;
    g.count = s - first;
    const double z0 = 0.5 * ((double)p.zab[first].x + p.zab[first].y);
    const double zl = 0.5 * ((double)p.zab[s - 1].x + p.zab[s - 1].y);
    const double h = g.count > 1 ? (zl - z0) / (g.count - 1) : 1.0;
    double dlo = 1e30, dhi = -1e30, devz = 0;
    for (int q = first; q < s; q++) {
      const double zm = 0.5 * ((double)p.zab[q].x + p.zab[q].y);
      devz = std::max(devz, std::fabs(zm - (z0 + (q - first) * h)));
      const double d = (double)p.zab[q].y - p.zab[q].x;
      dlo = std::min(dlo, d);
      dhi = std::max(dhi, d);
    }
    g.z0 = (float)z0;
    g.h = (float)h;
    const int mid = first + g.count / 2;
    g.dlo = p.zab[mid].y - p.zab[mid].x;  // typical axial slope of the run
    g.dhi = (float)(1.0 / h);             // 1 / h for the start estimate
    g.margin = (float)(devz + std::max(std::fabs(dlo), std::fabs(dhi)));
    g.dmin = (float)dlo;
    g.dmax = (float)dhi;
    g.devz = (float)devz;
    segs.push_back(g);
    first = s;
  }
  ok = ok && !segs.empty();
}

}  // namespace apirl

extern "C" void apirl_sino_grid(int nang, int nr, float rfov, float* ang_deg, float* r_mm) {
  // Sinogram2DinCylindrical3Dpet.cpp:46-60 (maxAng_deg is the int 180)
  float rinc = (2 * rfov) / nr;
  float phinc = (float)180 / nang;
  for (int i = 0; i < nang; i++) ang_deg[i] = i * phinc;
  for (int j = 0; j < nr; j++) r_mm[j] = rinc / 2 + j * rinc - rfov;
}

extern "C" int apirl_ring_table(int nrings, float zfov, int nseg, const int* nsino,
                                const int* minrd, const int* maxrd, float* z1m, float* z2m) {
  // Ring grid Sinogram3D.cpp:358-367; ring-pair lists :396-464; mean of the first and
  // last pair of each list, SiddonProjector.cpp:357-360.
  std::vector<float> zr(nrings);
  float zinc = (float)zfov / nrings;
  for (int i = 0; i < nrings; i++) zr[i] = zinc / 2 + i * zinc;
  int total = 0;
  for (int s = 0; s < nseg; s++) {
    int count = 0;
    for (int z1 = 0; z1 < 2 * nrings; z1++) {
      int n = 0, f1 = 0, f2 = 0, l1 = 0, l2 = 0;
      int z1a = z1;
      for (int z2 = 0; z2 < nrings; z2++) {
        if ((z1a - z2) <= maxrd[s] && (z1a - z2) >= minrd[s] && z1a >= 0 && z2 >= 0 &&
            z1a < nrings && z2 < nrings) {
          if (n == 0) { f1 = z1a; f2 = z2; }
          l1 = z1a; l2 = z2;
          n++;
        }
        z1a--;
      }
      if (n > 0) {
        z1m[total] = (zr[f1] + zr[l1]) / 2;
        z2m[total] = (zr[f2] + zr[l2]) / 2;
        total++;
        count++;
      }
    }
    if (count != nsino[s]) return -1;
  }
  return total;
}

extern "C" int apirl_subset_num_angles(int nang, int nsubsets, int subset) {
  return apirl::subset_num_angles(nang, nsubsets, subset);
}

namespace apirl {

// Entry crossings of every LOR (Siddon.cpp:322,330 / :626,637: alpha_x += alpha_x_u):
// for ray (a, r), out[(a nr + r) (nx + ny) + i] = the value after t float additions
// of axu to ax, where column i = i0 + (t + 1) iinc is the one that crossing enters,
// for every column of the image in the ray's direction; likewise rows after nx.
// Float addition here rounds as on the device (IEEE binary32, no contraction).
void build_xy_crossings(const apirl_image_desc& img, const std::vector<RayXY>& rays,
                        std::vector<float>& out) {
  const int nx = img.nx, ny = img.ny;
  out.assign(rays.size() * (size_t)(nx + ny), 0.f);
  for (size_t q = 0; q < rays.size(); q++) {
    const RayXY& t = rays[q];
    float* row = out.data() + q * (size_t)(nx + ny);
    const int iinc = ray_iinc(t.flags), jinc = ray_jinc(t.flags);
    // walk until past the far side of the image (the entry voxel may lie outside it)
    if (iinc != 0) {
      float a = t.ax;
      for (int i = t.i0 + iinc; iinc > 0 ? i < nx : i >= 0; i += iinc) {
        if (i >= 0 && i < nx) row[i] = a;
        a = a + t.axu;
      }
    }
    if (jinc != 0) {
      float a = t.ay;
      for (int j = t.j0 + jinc; jinc > 0 ? j < ny : j >= 0; j += jinc) {
        if (j >= 0 && j < ny) row[nx + j] = a;
        a = a + t.ayu;
      }
    }
  }
}

void build_f64s_rays(const Geo& g, const std::vector<RayXY>& rays, std::vector<RayXY>& out) {
  out = rays;
  for (RayXY& o : out) {
    SetupXY64 xy;
    setup_xy_f64(g, o.p0x, o.p0y, o.vx, o.vy, xy);
    o.amin = (float)xy.amin;
    o.amax = (float)xy.amax;
    o.ax = (float)xy.ax;
    o.ay = (float)xy.ay;
    o.axu = (float)xy.axu;
    o.ayu = (float)xy.ayu;
    o.i0 = xy.i0;
    o.j0 = xy.j0;
    o.nxy = std::abs(xy.i1 - xy.i0) + std::abs(xy.j1 - xy.j0);
  }
}

}  // namespace apirl
