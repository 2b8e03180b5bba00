/*
 * HipSiddonProjector -- the reference-side binding of the MI355X Siddon path.
 *
 * A Projector (recon/inc/Projector.h:48-85) whose Project / Backproject /
 * DivideAndBackproject run on a gfx950 GPU through the C-ABI of include/apirl_hip.h.
 * A maintainer adds this file and HipSiddonProjector.cpp to recon/ and selects it by
 * name where cmd/src/MLEM.cpp:305-346, cmd/src/OSEM.cpp:311-352 and
 * cmd/src/project.cpp:248-279 instantiate projectors.  It replaces the reference's
 * CuProjectorInterface (reconGPU/src/CuProjectorInterface.cu), which re-initialises
 * the device and re-uploads the geometry on every call; here the geometry plan of
 * each (image, sinogram) pair is built once and kept, and only the bins / voxels
 * move per call.
 *
 * Semantics follow SiddonProjector (recon/src/SiddonProjector.cpp): Project
 * overwrites, Backproject and DivideAndBackproject accumulate into the caller's
 * image.  Errors return false with getLastError() filled, as the reference does.
 * Each bin is the LOR of its sinogram's mean ring pair (SiddonProjector.cpp:349-361);
 * with setMultipleLorsPerBin(true) (Projector.h:54-63) the 3D calls return false.
 */
#ifndef HIP_SIDDON_PROJECTOR_H
#define HIP_SIDDON_PROJECTOR_H

#include <Images.h>
#include <Projector.h>
#include <Sinogram2D.h>
#include <Sinogram3D.h>

#include <map>
#include <string>

#include "apirl_hip.h"

class HipSiddonProjector : virtual Projector {
 public:
  /* gpuId: device index; backprojectionMode: APIRL_BP_GATHER (deterministic, the
   * default; Backproject and DivideAndBackproject alike), APIRL_BP_TILE (3D only,
   * deterministic, the fastest) or APIRL_BP_SCATTER (fp32 atomics). */
  HipSiddonProjector(int gpuId = 0, int backprojectionMode = APIRL_BP_GATHER);
  /* SiddonProjector(nSamplesOnDetector, nAxialSamplesOnDetector)
   * (recon/src/SiddonProjector.cpp:28-32): 3D bins of a cylindrical scanner traced as
   * numSamples x numAxialSamples LORs (apirl_plan_set_oversampling), in either
   * backprojection mode (the gather takes a bin's LORs in turn). */
  HipSiddonProjector(int gpuId, int backprojectionMode, int numSamples, int numAxialSamples);
  virtual ~HipSiddonProjector();

  bool Project(Image* image, Sinogram2D* projection);
  bool Backproject(Sinogram2D* projection, Image* image);
  bool DivideAndBackproject(Sinogram2D* measured, Sinogram2D* estimated, Image* image);
  bool Project(Image* image, Sinogram3D* projection);
  bool Backproject(Sinogram3D* projection, Image* image);
  bool DivideAndBackproject(Sinogram3D* measured, Sinogram3D* estimated, Image* image);

  std::string getLastError() { return strError; }

 private:
  apirl_plan* planFor(Image* image, Sinogram2D* sino);
  apirl_plan* planFor(Image* image, Sinogram3D* sino);
  bool ensure(float** buf, size_t* cap, size_t n);
  bool fail(const std::string& what);
  void copySinogramIn(Sinogram3D* s, float* host);
  void copySinogramOut(Sinogram3D* s, const float* host);

  bool multipleLorsRefused();

  apirl_ctx* ctx;
  int mode;
  int nsamples, naxial;  // LORs per 3D bin (transaxial, axial)
  std::string strError;
  std::map<std::string, apirl_plan*> plans;
  float *d_img, *d_sino, *d_sino2;
  size_t cap_img, cap_sino, cap_sino2;
};

#endif
