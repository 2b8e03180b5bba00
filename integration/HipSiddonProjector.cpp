/*
 * HipSiddonProjector.cpp -- see HipSiddonProjector.h.
 */
#include "HipSiddonProjector.h"

#include <Sinogram2DinCylindrical3Dpet.h>
#include <Sinogram2DinSiemensMmr.h>
#include <Sinogram3DCylindricalPet.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

void append(std::string& key, const void* p, size_t n) {
  key.append(static_cast<const char*>(p), n);
}

size_t bins3d(Sinogram3D* s) {
  size_t n = 0;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2D* s2 = s->getSegment(i)->getSinogram2D(j);
      n += (size_t)s2->getNumProj() * s2->getNumR();
    }
  return n;
}

}  // namespace

HipSiddonProjector::HipSiddonProjector(int gpuId, int backprojectionMode)
    : HipSiddonProjector(gpuId, backprojectionMode, 1, 1) {}

HipSiddonProjector::HipSiddonProjector(int gpuId, int backprojectionMode, int numSamples,
                                       int numAxialSamples)
    : ctx(nullptr), mode(backprojectionMode), nsamples(numSamples), naxial(numAxialSamples),
      d_img(nullptr), d_sino(nullptr), d_sino2(nullptr), cap_img(0), cap_sino(0),
      cap_sino2(0) {
  if (apirl_ctx_create(gpuId, &ctx) != APIRL_OK) {
    ctx = nullptr;
    strError = "HipSiddonProjector: no usable GPU for index " + std::to_string(gpuId);
  }
}

bool HipSiddonProjector::multipleLorsRefused() {
  // SiddonProjector.cpp:349-352: with useMultipleLorsPerBin every ring pair of a
  // sinogram is traced; this projector traces the mean pair only, so it says so
  // instead of returning another result
  if (!useMultipleLorsPerBin) return false;
  fail("HipSiddonProjector: useMultipleLorsPerBin (one LOR per ring pair) is not "
       "supported; each bin is traced on its sinogram's mean ring pair");
  return true;
}

HipSiddonProjector::~HipSiddonProjector() {
  if (!ctx) return;
  for (auto& kv : plans) apirl_plan_destroy(kv.second);
  if (d_img) apirl_free(ctx, d_img);
  if (d_sino) apirl_free(ctx, d_sino);
  if (d_sino2) apirl_free(ctx, d_sino2);
  apirl_ctx_destroy(ctx);
}

bool HipSiddonProjector::fail(const std::string& what) {
  // The reference's EM drivers ignore the projector's return value
  // (Mlem2d.cpp:151,167,275), so the message is also printed, as the reference
  // prints its own errors.
  strError = what;
  if (ctx) {
    const char* m = apirl_last_error(ctx);
    if (m && *m) strError += std::string(": ") + m;
  }
  fprintf(stderr, "%s\n", strError.c_str());
  return false;
}

bool HipSiddonProjector::ensure(float** buf, size_t* cap, size_t n) {
  if (*cap >= n) return true;
  if (*buf) apirl_free(ctx, *buf);
  *buf = nullptr;
  *cap = 0;
  void* p = nullptr;
  if (apirl_malloc(ctx, &p, n * sizeof(float)) != APIRL_OK) return false;
  *buf = static_cast<float*>(p);
  *cap = n;
  return true;
}

static apirl_image_desc image_desc(Image* image) {
  SizeImage s = image->getSize();
  apirl_image_desc d;
  d.nx = s.nPixelsX;
  d.ny = s.nPixelsY;
  d.nz = s.nDimensions == 2 ? 1 : s.nPixelsZ;
  d.dx = s.sizePixelX_mm;
  d.dy = s.sizePixelY_mm;
  d.dz = s.sizePixelZ_mm;
  return d;
}

apirl_plan* HipSiddonProjector::planFor(Image* image, Sinogram2D* sino) {
  // A 2D sinogram of a cylindrical scanner: LORs from Sinogram2DinCylindrical3Dpet
  // (Sinogram2DinCylindrical3Dpet.cpp:249-266).
  Sinogram2DinCylindrical3Dpet* cyl = dynamic_cast<Sinogram2DinCylindrical3Dpet*>(sino);
  if (!cyl) return nullptr;
  apirl_image_desc im = image_desc(image);
  apirl_sino_desc sd;
  memset(&sd, 0, sizeof(sd));
  sd.dims = 2;
  sd.nang = sino->getNumProj();
  sd.nr = sino->getNumR();
  sd.nsino = 1;
  sd.rfov_mm = sino->getRadioFov_mm();
  sd.rscanner_mm = cyl->getRadioScanner_mm();
  sd.ang_deg = sino->getAngPtr();
  sd.r_mm = sino->getRPtr();
  std::string key = "2";
  append(key, &im, sizeof(im));
  append(key, &sd.nang, 3 * sizeof(int));
  append(key, &sd.rfov_mm, 2 * sizeof(float));
  append(key, sd.ang_deg, sd.nang * sizeof(float));
  append(key, sd.r_mm, sd.nr * sizeof(float));
  auto it = plans.find(key);
  if (it != plans.end()) return it->second;
  apirl_plan* p = nullptr;
  if (apirl_plan_create(ctx, &im, &sd, &p) != APIRL_OK) return nullptr;
  plans[key] = p;
  return p;
}

apirl_plan* HipSiddonProjector::planFor(Image* image, Sinogram3D* sino) {
  // Every 2D sinogram of a Sinogram3D shares the angle / r grids; each contributes the
  // mean of its first and last ring pair (SiddonProjector.cpp:357-360).
  const int nseg = sino->getNumSegments();
  if (nseg < 1 || sino->getSegment(0)->getNumSinograms() < 1) return nullptr;
  Sinogram2DinCylindrical3Dpet* s00 = sino->getSegment(0)->getSinogram2D(0);
  std::vector<float> z1, z2;
  for (int i = 0; i < nseg; i++)
    for (int j = 0; j < sino->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2DinCylindrical3Dpet* s2 = sino->getSegment(i)->getSinogram2D(j);
      const int nz = s2->getNumZ();
      z1.push_back((s2->getAxialValue1FromList(0) + s2->getAxialValue1FromList(nz - 1)) / 2);
      z2.push_back((s2->getAxialValue2FromList(0) + s2->getAxialValue2FromList(nz - 1)) / 2);
    }
  // The LORs are those of the 2D sinograms' own class (getPointsFromLor is virtual):
  // the Siemens mMR's (Sinogram2DinSiemensMmr.cpp:175-192), whose grids the library
  // builds from the scanner's constants, or the cylindrical scanner's from the
  // sinogram's angle and radial grids.  An mMR sinogram whose angles are not the full
  // grid a * 180 / nang (a subset) is not supported.
  const bool mmr = dynamic_cast<Sinogram2DinSiemensMmr*>(s00) != nullptr;
  if (mmr) {
    const float* ang = s00->getAngPtr();
    const float inc = (float)180.0f / s00->getNumProj();
    for (int a = 0; a < s00->getNumProj(); a++)
      if (ang[a] != a * inc) return nullptr;
  }
  apirl_image_desc im = image_desc(image);
  apirl_sino_desc sd;
  memset(&sd, 0, sizeof(sd));
  sd.dims = 3;
  sd.nang = s00->getNumProj();
  sd.nr = s00->getNumR();
  sd.nsino = (int)z1.size();
  sd.rfov_mm = s00->getRadioFov_mm();
  sd.rscanner_mm = s00->getRadioScanner_mm();
  sd.ring_width_mm = sino->getAxialFoV_mm() / sino->getNumRings();  // SiddonProjector.cpp:335
  sd.ang_deg = mmr ? nullptr : s00->getAngPtr();
  sd.r_mm = mmr ? nullptr : s00->getRPtr();
  sd.z1_mean_mm = z1.data();
  sd.z2_mean_mm = z2.data();
  if (mmr && nsamples * naxial != 1) return nullptr;  // defined for the cylindrical scanner
  std::string key = mmr ? "3m" : "3";
  append(key, &im, sizeof(im));
  append(key, &sd.nang, 3 * sizeof(int));
  append(key, &sd.rfov_mm, 3 * sizeof(float));
  if (!mmr) {
    append(key, sd.ang_deg, sd.nang * sizeof(float));
    append(key, sd.r_mm, sd.nr * sizeof(float));
  }
  append(key, z1.data(), z1.size() * sizeof(float));
  append(key, z2.data(), z2.size() * sizeof(float));
  auto it = plans.find(key);
  if (it != plans.end()) return it->second;
  apirl_plan* p = nullptr;
  if (apirl_plan_create(ctx, &im, &sd, &p) != APIRL_OK) return nullptr;
  if (mmr && apirl_plan_set_scanner(p, APIRL_SCANNER_MMR) != APIRL_OK) {
    apirl_plan_destroy(p);
    return nullptr;
  }
  if (nsamples * naxial != 1 && apirl_plan_set_oversampling(p, nsamples, naxial) != APIRL_OK) {
    apirl_plan_destroy(p);
    return nullptr;
  }
  plans[key] = p;
  return p;
}

void HipSiddonProjector::copySinogramIn(Sinogram3D* s, float* host) {
  size_t off = 0;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2D* s2 = s->getSegment(i)->getSinogram2D(j);
      const size_t n = (size_t)s2->getNumProj() * s2->getNumR();
      memcpy(host + off, s2->getSinogramPtr(), n * sizeof(float));
      off += n;
    }
}

void HipSiddonProjector::copySinogramOut(Sinogram3D* s, const float* host) {
  size_t off = 0;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2D* s2 = s->getSegment(i)->getSinogram2D(j);
      const size_t n = (size_t)s2->getNumProj() * s2->getNumR();
      memcpy(s2->getSinogramPtr(), host + off, n * sizeof(float));
      off += n;
    }
}

// ---------------------------------------------------------------- 2D

bool HipSiddonProjector::Project(Image* image, Sinogram2D* projection) {
  if (!ctx) return false;
  apirl_plan* p = planFor(image, projection);
  if (!p) return fail("HipSiddonProjector::Project: unsupported 2D geometry");
  const size_t nv = apirl_plan_image_voxels(p), nb = apirl_plan_subset_bins(p, 0, 1);
  if (!ensure(&d_img, &cap_img, nv) || !ensure(&d_sino, &cap_sino, nb))
    return fail("HipSiddonProjector::Project: out of device memory");
  // bins whose LOR the reference rejects keep their value (SiddonProjector.cpp:166)
  if (apirl_memcpy_h2d(ctx, d_img, image->getPixelsPtr(), nv * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino, projection->getSinogramPtr(), nb * 4) != APIRL_OK ||
      apirl_project(p, d_img, d_sino, 0, 1, APIRL_PREC_REFERENCE) != APIRL_OK ||
      apirl_memcpy_d2h(ctx, projection->getSinogramPtr(), d_sino, nb * 4) != APIRL_OK)
    return fail("HipSiddonProjector::Project");
  return true;
}

bool HipSiddonProjector::Backproject(Sinogram2D* projection, Image* image) {
  if (!ctx) return false;
  apirl_plan* p = planFor(image, projection);
  if (!p) return fail("HipSiddonProjector::Backproject: unsupported 2D geometry");
  const size_t nv = apirl_plan_image_voxels(p), nb = apirl_plan_subset_bins(p, 0, 1);
  if (!ensure(&d_img, &cap_img, nv) || !ensure(&d_sino, &cap_sino, nb))
    return fail("HipSiddonProjector::Backproject: out of device memory");
  if (apirl_memcpy_h2d(ctx, d_img, image->getPixelsPtr(), nv * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino, projection->getSinogramPtr(), nb * 4) != APIRL_OK ||
      apirl_backproject(p, d_sino, d_img, 0, 1, mode, APIRL_PREC_REFERENCE) != APIRL_OK ||
      apirl_memcpy_d2h(ctx, image->getPixelsPtr(), d_img, nv * 4) != APIRL_OK)
    return fail("HipSiddonProjector::Backproject");
  return true;
}

bool HipSiddonProjector::DivideAndBackproject(Sinogram2D* measured, Sinogram2D* estimated,
                                              Image* image) {
  if (!ctx) return false;
  apirl_plan* p = planFor(image, measured);
  if (!p) return fail("HipSiddonProjector::DivideAndBackproject: unsupported 2D geometry");
  const size_t nv = apirl_plan_image_voxels(p), nb = apirl_plan_subset_bins(p, 0, 1);
  if (!ensure(&d_img, &cap_img, nv) || !ensure(&d_sino, &cap_sino, nb) ||
      !ensure(&d_sino2, &cap_sino2, nb))
    return fail("HipSiddonProjector::DivideAndBackproject: out of device memory");
  if (apirl_memcpy_h2d(ctx, d_img, image->getPixelsPtr(), nv * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino, measured->getSinogramPtr(), nb * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino2, estimated->getSinogramPtr(), nb * 4) != APIRL_OK ||
      apirl_divide_backproject(p, d_sino, d_sino2, d_img, 0, 1, mode) != APIRL_OK ||
      apirl_memcpy_d2h(ctx, image->getPixelsPtr(), d_img, nv * 4) != APIRL_OK)
    return fail("HipSiddonProjector::DivideAndBackproject");
  return true;
}

// ---------------------------------------------------------------- 3D

bool HipSiddonProjector::Project(Image* image, Sinogram3D* projection) {
  if (!ctx) return false;
  if (multipleLorsRefused()) return false;
  apirl_plan* p = planFor(image, projection);
  if (!p) return fail("HipSiddonProjector::Project: unsupported 3D geometry");
  const size_t nv = apirl_plan_image_voxels(p), nb = bins3d(projection);
  if (nb != apirl_plan_subset_bins(p, 0, 1)) return fail("HipSiddonProjector: bin count");
  if (!ensure(&d_img, &cap_img, nv) || !ensure(&d_sino, &cap_sino, nb))
    return fail("HipSiddonProjector::Project: out of device memory");
  std::vector<float> host(nb);
  if (apirl_memcpy_h2d(ctx, d_img, image->getPixelsPtr(), nv * 4) != APIRL_OK ||
      apirl_project(p, d_img, d_sino, 0, 1, APIRL_PREC_REFERENCE) != APIRL_OK ||
      apirl_memcpy_d2h(ctx, host.data(), d_sino, nb * 4) != APIRL_OK)
    return fail("HipSiddonProjector::Project");
  copySinogramOut(projection, host.data());
  return true;
}

bool HipSiddonProjector::Backproject(Sinogram3D* projection, Image* image) {
  if (!ctx) return false;
  if (multipleLorsRefused()) return false;
  apirl_plan* p = planFor(image, projection);
  if (!p) return fail("HipSiddonProjector::Backproject: unsupported 3D geometry");
  const size_t nv = apirl_plan_image_voxels(p), nb = bins3d(projection);
  if (nb != apirl_plan_subset_bins(p, 0, 1)) return fail("HipSiddonProjector: bin count");
  if (!ensure(&d_img, &cap_img, nv) || !ensure(&d_sino, &cap_sino, nb))
    return fail("HipSiddonProjector::Backproject: out of device memory");
  std::vector<float> host(nb);
  copySinogramIn(projection, host.data());
  if (apirl_memcpy_h2d(ctx, d_img, image->getPixelsPtr(), nv * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino, host.data(), nb * 4) != APIRL_OK ||
      apirl_backproject(p, d_sino, d_img, 0, 1, mode, APIRL_PREC_REFERENCE) != APIRL_OK ||
      apirl_memcpy_d2h(ctx, image->getPixelsPtr(), d_img, nv * 4) != APIRL_OK)
    return fail("HipSiddonProjector::Backproject");
  return true;
}

bool HipSiddonProjector::DivideAndBackproject(Sinogram3D* measured, Sinogram3D* estimated,
                                              Image* image) {
  if (!ctx) return false;
  if (multipleLorsRefused()) return false;
  apirl_plan* p = planFor(image, measured);
  if (!p) return fail("HipSiddonProjector::DivideAndBackproject: unsupported 3D geometry");
  const size_t nv = apirl_plan_image_voxels(p), nb = bins3d(measured);
  if (nb != apirl_plan_subset_bins(p, 0, 1) || nb != bins3d(estimated))
    return fail("HipSiddonProjector: bin count");
  if (!ensure(&d_img, &cap_img, nv) || !ensure(&d_sino, &cap_sino, nb) ||
      !ensure(&d_sino2, &cap_sino2, nb))
    return fail("HipSiddonProjector::DivideAndBackproject: out of device memory");
  std::vector<float> hm(nb), he(nb);
  copySinogramIn(measured, hm.data());
  copySinogramIn(estimated, he.data());
  if (apirl_memcpy_h2d(ctx, d_img, image->getPixelsPtr(), nv * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino, hm.data(), nb * 4) != APIRL_OK ||
      apirl_memcpy_h2d(ctx, d_sino2, he.data(), nb * 4) != APIRL_OK ||
      apirl_divide_backproject(p, d_sino, d_sino2, d_img, 0, 1, mode) != APIRL_OK ||
      apirl_memcpy_d2h(ctx, image->getPixelsPtr(), d_img, nv * 4) != APIRL_OK)
    return fail("HipSiddonProjector::DivideAndBackproject");
  return true;
}
