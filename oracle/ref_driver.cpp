// Test-infrastructure driver around the reference's own compiled CPU Siddon path
// (oracle/_ref/libapirl_ref.so, built by oracle/ref.mk from the reference's
// sources).  It is a checker and a CPU baseline only: nothing in the product
// links it.  It constructs the reference's Image / Sinogram objects the way its
// command-line tools do and calls the reference's SiddonProjector, Mlem2d,
// MlemSinogram3d and OsemSinogram3d, then writes the results as raw float32.
//
// Usage:  ref_driver <command> key=value ...   (all arrays raw little-endian f32)
//   siddon2d  nx dx nproj nr rfov rscanner out          per-bin ray lists
//   siddon3d  <geom3d> out                              per-bin ray lists
//   proj2d    nx dx nproj nr rfov rscanner img out
//   back2d    nx dx nproj nr rfov rscanner sino out
//   mlem2d    nx dx nproj nr rscanner iters meas init outdir
//   divback2d nx dx nproj nr rfov rscanner meas est out
//   mlem2dmulti nx ny nz dx dz header rfov zfov rscanner iters outdir out
//   divback3d <geom3d> meas est out
//   proj3d    <geom3d> img out
//   back3d    <geom3d> sino out
//   osem3d    <geom3d> subsets iters meas init outdir [add mult guard]
//   mlem3d    <geom3d> iters meas init outdir
//   rays3d    <geom3d> img start stride out             sampled reference rays
//   osem3d_sub <geom3d> subsets subset meas             one timed OSEM sub-iteration
//   img_write / img_read / sino3d_write / sino3d_read / sino2d_write / sino2d_read
//                                                       the on-disk formats
//   par_read file method key / itf_read header key     one key lookup: "rc<TAB>value"
//   log_row   n in out                                 Logger::writeRowOfNumbers
//   mmr3d     header nx ny nz dx dz img out_proj out_back   mMR projection / backprojection
//   mmr_mlem3d header nx ny nz dx dz init iters outdir [add mult]  MlemSinogram3d, mMR
//   mmr_osem3d header nx ny nz dx dz init iters subsets outdir [add mult]
//                                                       OsemSinogram3d on an mMR sinogram
//   mmr_subset_proj header nx ny nz dx dz img subset subsets out   projection of a subset
//   <geom3d> = nx ny nz dx dz nproj nr rings rfov zfov rscanner span mrd
// The reference's 2D projector checks only lower bounds and its 3D projector only
// upper bounds (SiddonProjector.cpp:178,417), so an entry voxel index of -1 or n can
// read outside the pixel array (undefined behaviour; it segfaults when the array is
// page-aligned).  For the calls this driver makes directly, the pixel array is moved
// into a buffer with zeroed guard zones, so such a read returns 0.  That swap needs
// Image's private pixel pointer: oracle/ref.mk compiles this driver with
// -fno-access-control, leaving the reference's headers as they are.
#include <Images.h>
#include <Sinogram2DinCylindrical3Dpet.h>
#include <Sinogram3DCylindricalPet.h>
#include <Sinogram3DSiemensMmr.h>
#include <SiddonProjector.h>
#include <Siddon.h>
#include <Mlem2d.h>
#include <Mlem2dMultiple.h>
#include <MlemSinogram3d.h>
#include <OsemSinogram3d.h>
#include <omp.h>

#ifdef WITH_HIP
#include "HipSiddonProjector.h"
#endif

#include <Interfile.h>
#include <Logger.h>
#include <ParametersFile.h>
#include <malloc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

namespace {

std::map<std::string, std::string> g_args;

const char* arg(const char* k) {
  auto it = g_args.find(k);
  if (it == g_args.end()) {
    fprintf(stderr, "ref_driver: missing argument %s\n", k);
    exit(2);
  }
  return it->second.c_str();
}
int iarg(const char* k) { return atoi(arg(k)); }
bool has_arg(const char* k) { return g_args.count(k) != 0; }
float farg(const char* k) { return strtof(arg(k), nullptr); }

std::vector<float> read_f32(const char* path, size_t n) {
  std::vector<float> v(n);
  FILE* f = fopen(path, "rb");
  if (!f || fread(v.data(), 4, n, f) != n) {
    fprintf(stderr, "ref_driver: cannot read %zu floats from %s\n", n, path);
    exit(3);
  }
  fclose(f);
  return v;
}

void write_f32(const char* path, const float* p, size_t n) {
  FILE* f = fopen(path, "wb");
  if (!f || fwrite(p, 4, n, f) != n) {
    fprintf(stderr, "ref_driver: cannot write %s\n", path);
    exit(4);
  }
  fclose(f);
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

SizeImage size2d() {
  SizeImage s;
  s.nDimensions = 2;
  s.nPixelsX = s.nPixelsY = iarg("nx");
  s.nPixelsZ = 1;
  s.sizePixelX_mm = s.sizePixelY_mm = farg("dx");
  s.sizePixelZ_mm = 1.0f;
  return s;
}

SizeImage size3d() {
  SizeImage s;
  s.nDimensions = 3;
  s.nPixelsX = iarg("nx");
  s.nPixelsY = iarg("ny");
  s.nPixelsZ = iarg("nz");
  s.sizePixelX_mm = s.sizePixelY_mm = farg("dx");
  s.sizePixelZ_mm = farg("dz");
  return s;
}

// Segment table for a span / maximum ring difference, ordered 0, +1, -1, +2, -2 ...
// (the order the .par / interfile segment lists use for cylindrical scanners).
struct SegTable {
  std::vector<int> nsino, minrd, maxrd;
};

SegTable seg_table(int rings, int span, int mrd) {
  SegTable t;
  int h = (span - 1) / 2;
  t.minrd.push_back(-h);
  t.maxrd.push_back(h);
  t.nsino.push_back(2 * rings - 1);
  for (int lo = h + 1; lo <= mrd; lo += span) {
    int hi = lo + span - 1;
    if (hi > mrd) hi = mrd;
    int n = 0;  // ring sums a+b reachable with lo <= a-b <= hi, 0 <= a,b < rings
    for (int sum = 0; sum <= 2 * rings - 2; sum++)
      for (int b = 0; b < rings; b++) {
        int a = sum - b;
        if (a >= 0 && a < rings && a - b >= lo && a - b <= hi) { n++; break; }
      }
    t.minrd.push_back(lo);  t.maxrd.push_back(hi);  t.nsino.push_back(n);
    t.minrd.push_back(-hi); t.maxrd.push_back(-lo); t.nsino.push_back(n);
  }
  return t;
}

// Builds a cylindrical 3D sinogram with the ring grid the reference's interfile
// reader computes (Sinogram3D.cpp:358-367) and its ring-pair lists.
Sinogram3DCylindricalPet* make_sino3d() {
  int rings = iarg("rings");
  float zfov = farg("zfov");
  SegTable t = seg_table(rings, iarg("span"), iarg("mrd"));
  // The reference's Sinogram3D(float, float) base constructor leaves ptrAngValues_deg
  // unset and the parametric constructor then realloc()s it
  // (Sinogram3DCylindricalPet.cpp:55); constructing on zeroed storage makes that a
  // realloc(NULL) instead of undefined behaviour.
  void* mem = calloc(1, sizeof(Sinogram3DCylindricalPet));
  Sinogram3DCylindricalPet* s = new (mem) Sinogram3DCylindricalPet(
      iarg("nproj"), iarg("nr"), rings, farg("rfov"), zfov, farg("rscanner"), (int)t.nsino.size(),
      t.nsino.data(), t.minrd.data(), t.maxrd.data());
  float* z = s->getAxialPtr();
  float zIncrement = (float)zfov / rings;
  for (int i = 0; i < rings; i++) z[i] = zIncrement / 2 + i * zIncrement;
  if (!s->initRingConfig(t.nsino.data())) {
    fprintf(stderr, "ref_driver: ring configuration mismatch\n");
    exit(5);
  }
  return s;
}

size_t bins3d(Sinogram3D* s) {
  size_t n = 0;
  for (int i = 0; i < s->getNumSegments(); i++)
    n += (size_t)s->getSegment(i)->getNumSinograms() * s->getNumProj() * s->getNumR();
  return n;
}

void sino3d_from(Sinogram3D* s, const float* src) {
  size_t off = 0;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2D* s2 = s->getSegment(i)->getSinogram2D(j);
      size_t n = (size_t)s2->getNumProj() * s2->getNumR();
      memcpy(s2->getSinogramPtr(), src + off, n * 4);
      off += n;
    }
}

std::vector<float> sino3d_to(Sinogram3D* s) {
  std::vector<float> v(bins3d(s));
  size_t off = 0;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2D* s2 = s->getSegment(i)->getSinogram2D(j);
      size_t n = (size_t)s2->getNumProj() * s2->getNumR();
      memcpy(v.data() + off, s2->getSinogramPtr(), n * 4);
      off += n;
    }
  return v;
}

// Exposes the protected results of the reference's EM drivers.
struct Mlem2dProbe : Mlem2d {
  using Mlem2d::Mlem2d;
  float* lik() { return likelihoodValues; }
  Image* sens() { return sensitivityImage; }
};
// corrections(): the additive / multiplicative sinograms set as the reference's
// setAdditiveProjection / setMultiplicativeProjection leave them
// (MlemSinogram3d.cpp:17-42), but from sinograms built here: those loaders copy the
// input sinogram, and the cylindrical copy constructor leaves the scanner radius
// unset, which getSubset then passes on to the multiplicative subsets.
struct Mlem3dProbe : MlemSinogram3d {
  using MlemSinogram3d::MlemSinogram3d;
  float* lik() { return likelihoodValues; }
  Image* sens() { return sensitivityImage; }
  void corrections(Sinogram3D* add, Sinogram3D* mult) {
    if (add) { additiveProjection = add; enableAdditiveTerm = true; }
    if (mult) { multiplicativeProjection = mult; enableMultiplicativeTerm = true; }
  }
};
struct OsemProbe : OsemSinogram3d {
  using OsemSinogram3d::OsemSinogram3d;
  float* lik() { return likelihoodValues; }
  Image* sens(int s) { return sensitivityImages[s]; }
  void corrections(Sinogram3D* add, Sinogram3D* mult) {
    if (add) { additiveProjection = add; enableAdditiveTerm = true; }
    if (mult) { multiplicativeProjection = mult; enableMultiplicativeTerm = true; }
  }
};


// Swaps an Image's pixel array for a copy inside zero guard zones of one slice plus
// one row on each side; restores the original array (owned by Image) on destruction.
struct GuardedPixels {
  Image* img;
  float* orig;
  float* buf;
  size_t guard, n;
  explicit GuardedPixels(Image* im) : img(im) {
    SizeImage s = im->getSize();
    n = (size_t)im->getPixelCount();
    guard = (size_t)s.nPixelsX * s.nPixelsY + s.nPixelsX;
    buf = (float*)calloc(n + 2 * guard, sizeof(float));
    orig = im->pixels;
    memcpy(buf + guard, orig, n * 4);
    im->pixels = buf + guard;
  }
  ~GuardedPixels() {
    memcpy(orig, img->pixels, n * 4);
    img->pixels = orig;
    free(buf);
  }
};

// The reference's SiddonProjector with every image it reads or writes held inside
// GuardedPixels for the call, so the 3D projector's out-of-range reads before the
// pixel array (see above) return 0 also inside the reference's EM drivers, which
// allocate their images themselves.
class GuardedSiddon : public SiddonProjector {
 public:
  using SiddonProjector::SiddonProjector;
  bool Project(Image* image, Sinogram3D* projection) {
    GuardedPixels g(image);
    return SiddonProjector::Project(image, projection);
  }
  bool Backproject(Sinogram3D* sino, Image* image) {
    GuardedPixels g(image);
    return SiddonProjector::Backproject(sino, image);
  }
  bool DivideAndBackproject(Sinogram3D* meas, Sinogram3D* est, Image* image) {
    GuardedPixels g(image);
    return SiddonProjector::DivideAndBackproject(meas, est, image);
  }
};

void report_time(const char* what, double t0) {
  printf("{\"what\": \"%s\", \"seconds\": %.6f, \"threads\": %d}\n", what, now_s() - t0,
         omp_get_max_threads());
  fflush(stdout);
}

// projector=siddon (default): the reference's SiddonProjector.
// projector=hip: the MI355X path through its reference-side binding
// (integration/HipSiddonProjector.cpp); bp=gather|scatter picks the backprojector.
Projector* make_projector() {
  const char* which = has_arg("projector") ? arg("projector") : "siddon";
  if (strcmp(which, "hip") == 0) {
#ifdef WITH_HIP
    int mode = (has_arg("bp") && strcmp(arg("bp"), "scatter") == 0) ? APIRL_BP_SCATTER
                                                                    : APIRL_BP_GATHER;
    Projector* hp = (has_arg("ns") || has_arg("nas"))
                        ? (Projector*)new HipSiddonProjector(0, mode,
                                                             has_arg("ns") ? iarg("ns") : 1,
                                                             has_arg("nas") ? iarg("nas") : 1)
                        : (Projector*)new HipSiddonProjector(0, mode);
    // multilor=1: Projector::setMultipleLorsPerBin(true) (Projector.h:60)
    if (has_arg("multilor")) hp->setMultipleLorsPerBin(iarg("multilor") != 0);
    return hp;
#else
    fprintf(stderr, "ref_driver: built without the HIP projector\n");
    exit(7);
#endif
  }
  if (has_arg("ns") || has_arg("nas"))  // oversampled LORs (SiddonProjector.cpp:28-32)
    return (Projector*)new SiddonProjector(has_arg("ns") ? iarg("ns") : 1,
                                           has_arg("nas") ? iarg("nas") : 1);
  return (Projector*)new SiddonProjector();
}

void write_lists(const char* path, std::vector<int>& counts, std::vector<SiddonSegment>& segs) {
  FILE* f = fopen(path, "wb");
  int nb = (int)counts.size();
  fwrite(&nb, 4, 1, f);
  fwrite(counts.data(), 4, counts.size(), f);
  for (auto& s : segs) {
    fwrite(&s.IndexX, 4, 1, f);
    fwrite(&s.IndexY, 4, 1, f);
    fwrite(&s.IndexZ, 4, 1, f);
    fwrite(&s.Segment, 4, 1, f);
  }
  fclose(f);
}

int cmd_siddon2d() {
  Image img(size2d());
  Sinogram2DinCylindrical3Dpet sino(iarg("nproj"), iarg("nr"), farg("rfov"), farg("rscanner"));
  std::vector<int> counts;
  std::vector<SiddonSegment> segs;
  for (int i = 0; i < sino.getNumProj(); i++)
    for (int j = 0; j < sino.getNumR(); j++) {
      Point2D p1, p2;
      float geom;
      int len = 0;
      if (sino.getPointsFromLor(i, j, &p1, &p2, &geom)) {
        Line2D lor;
        lor.P0 = p1;
        lor.Vx = p2.X - p1.X;
        lor.Vy = p2.Y - p1.Y;
        SiddonSegment* w = nullptr;
        Siddon(lor, &img, &w, &len, 1);
        for (int l = 0; l < len; l++) { w[l].IndexZ = 0; segs.push_back(w[l]); }
        if (len > 0) free(w);
      } else {
        len = -1;
      }
      counts.push_back(len);
    }
  write_lists(arg("out"), counts, segs);
  return 0;
}

int cmd_siddon3d() {
  Image img(size3d());
  Sinogram3DCylindricalPet* s = make_sino3d();
  float w = s->getAxialFoV_mm() / s->getNumRings();
  std::vector<int> counts;
  std::vector<SiddonSegment> segs;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2DinCylindrical3Dpet* s2 = s->getSegment(i)->getSinogram2D(j);
      float z1 = (s2->getAxialValue1FromList(0) + s2->getAxialValue1FromList(s2->getNumZ() - 1)) / 2;
      float z2 = (s2->getAxialValue2FromList(0) + s2->getAxialValue2FromList(s2->getNumZ() - 1)) / 2;
      for (int k = 0; k < s2->getNumProj(); k++)
        for (int l = 0; l < s2->getNumR(); l++) {
          Point3D p1, p2;
          float geom;
          s2->getPointsFromLor(k, l, 0, &p1, &p2, &geom);
          p1.Z = z1 - w / 2 + w / 2 + 0 * w;
          p2.Z = z2 - w / 2 + w / 2 + 0 * w;
          Line3D lor;
          lor.P0 = p1;
          lor.Vx = p2.X - p1.X;
          lor.Vy = p2.Y - p1.Y;
          lor.Vz = p2.Z - p1.Z;
          SiddonSegment* wl = nullptr;
          int len = 0;
          Siddon(lor, &img, &wl, &len, 1);
          for (int n = 0; n < len; n++) segs.push_back(wl[n]);
          if (len > 0) free(wl);
          counts.push_back(len);
        }
    }
  write_lists(arg("out"), counts, segs);
  return 0;
}

int cmd_proj2d() {
  Image img(size2d());
  size_t npix = (size_t)img.getPixelCount();
  std::vector<float> x = read_f32(arg("img"), npix);
  memcpy(img.getPixelsPtr(), x.data(), npix * 4);
  Sinogram2DinCylindrical3Dpet sino(iarg("nproj"), iarg("nr"), farg("rfov"), farg("rscanner"));
  Projector* pp = make_projector();
  double t0;
  {
  GuardedPixels guarded(&img);
  t0 = now_s();
  pp->Project(&img, &sino);
  report_time("proj2d", t0);
  }
  write_f32(arg("out"), sino.getSinogramPtr(), (size_t)sino.getNumProj() * sino.getNumR());
  return 0;
}

int cmd_back2d() {
  Image img(size2d());
  Sinogram2DinCylindrical3Dpet sino(iarg("nproj"), iarg("nr"), farg("rfov"), farg("rscanner"));
  size_t nb = (size_t)sino.getNumProj() * sino.getNumR();
  std::vector<float> y = read_f32(arg("sino"), nb);
  memcpy(sino.getSinogramPtr(), y.data(), nb * 4);
  Projector* pp = make_projector();
  double t0;
  {
  GuardedPixels guarded(&img);
  t0 = now_s();
  pp->Backproject(&sino, &img);
  report_time("back2d", t0);
  }
  write_f32(arg("out"), img.getPixelsPtr(), (size_t)img.getPixelCount());
  return 0;
}

// divback2d: Projector::DivideAndBackproject(meas, est, img) on a 2D sinogram.
int cmd_divback2d() {
  Image img(size2d());
  Sinogram2DinCylindrical3Dpet meas(iarg("nproj"), iarg("nr"), farg("rfov"), farg("rscanner"));
  Sinogram2DinCylindrical3Dpet est(iarg("nproj"), iarg("nr"), farg("rfov"), farg("rscanner"));
  size_t nb = (size_t)meas.getNumProj() * meas.getNumR();
  std::vector<float> m = read_f32(arg("meas"), nb), e = read_f32(arg("est"), nb);
  memcpy(meas.getSinogramPtr(), m.data(), nb * 4);
  memcpy(est.getSinogramPtr(), e.data(), nb * 4);
  Projector* pp = make_projector();
  {
  GuardedPixels guarded(&img);
  pp->DivideAndBackproject(&meas, &est, &img);
  }
  write_f32(arg("out"), img.getPixelsPtr(), (size_t)img.getPixelCount());
  return 0;
}

int cmd_mlem2d() {
  Image init(size2d());
  size_t npix = (size_t)init.getPixelCount();
  std::vector<float> x0 = read_f32(arg("init"), npix);
  memcpy(init.getPixelsPtr(), x0.data(), npix * 4);
  // The sinogram is built with the scanner radius as FOV; Mlem2d resets the radial
  // grid to the image FOV (Mlem2d.cpp:12) exactly as it does for the MLEM tool.
  Sinogram2DinCylindrical3Dpet sino(iarg("nproj"), iarg("nr"), farg("rscanner"), farg("rscanner"));
  size_t nb = (size_t)sino.getNumProj() * sino.getNumR();
  std::vector<float> y = read_f32(arg("meas"), nb);
  memcpy(sino.getSinogramPtr(), y.data(), nb * 4);
  Projector* pp = make_projector();
  std::string dir = arg("outdir");
  int iters = iarg("iters");
  Mlem2dProbe mlem(&sino, &init, dir, dir + "/mlem2d", iters, 0, false, false, pp, pp);
  double t0 = now_s();
  if (!mlem.Reconstruct()) return 6;
  report_time("mlem2d", t0);
  write_f32((dir + "/final.f32").c_str(), mlem.getReconstructedImage()->getPixelsPtr(), npix);
  write_f32((dir + "/sens.f32").c_str(), mlem.sens()->getPixelsPtr(), npix);
  write_f32((dir + "/lik.f32").c_str(), mlem.lik(), iters + 1);
  return 0;
}

// mlem2dmulti: Mlem2dMultiple (one Mlem2d per slice) on the multi-slice 2D sinogram
// file `header` read by Sinograms2DinCylindrical3Dpet(header, rfov, zfov, rscanner);
// the image grid is nx ny nz dx dz.  Writes the reconstructed volume to `out`.
int cmd_mlem2dmulti() {
  Sinograms2DinCylindrical3Dpet sinos(std::string(arg("header")), farg("rfov"), farg("zfov"),
                                      farg("rscanner"));
  Image init(size3d());
  init.fillConstant(1);
  Projector* pp = make_projector();
  std::string dir = arg("outdir");
  Mlem2dMultiple mlem(&sinos, &init, dir, dir + "/multi", iarg("iters"), 0, false, false, pp, pp);
  double t0 = now_s();
  if (!mlem.Reconstruct()) return 6;
  report_time("mlem2dmulti", t0);
  write_f32(arg("out"), mlem.getReconstructedImage()->getPixelsPtr(),
            (size_t)init.getPixelCount());
  return 0;
}

int cmd_proj3d() {
  Image img(size3d());
  size_t npix = (size_t)img.getPixelCount();
  std::vector<float> x = read_f32(arg("img"), npix);
  memcpy(img.getPixelsPtr(), x.data(), npix * 4);
  Sinogram3DCylindricalPet* s = make_sino3d();
  Projector* pp = make_projector();
  double t0;
  {
  GuardedPixels guarded(&img);
  t0 = now_s();
  pp->Project(&img, s);
  report_time("proj3d", t0);
  }
  std::vector<float> out = sino3d_to(s);
  write_f32(arg("out"), out.data(), out.size());
  return 0;
}

int cmd_back3d() {
  Image img(size3d());
  Sinogram3DCylindricalPet* s = make_sino3d();
  std::vector<float> y = read_f32(arg("sino"), bins3d(s));
  sino3d_from(s, y.data());
  Projector* pp = make_projector();
  double t0;
  {
  GuardedPixels guarded(&img);
  t0 = now_s();
  pp->Backproject(s, &img);
  report_time("back3d", t0);
  }
  write_f32(arg("out"), img.getPixelsPtr(), (size_t)img.getPixelCount());
  return 0;
}

// divback3d: Projector::DivideAndBackproject(meas, est, img) on a 3D sinogram.
int cmd_divback3d() {
  Image img(size3d());
  Sinogram3DCylindricalPet* m = make_sino3d();
  Sinogram3DCylindricalPet* e = make_sino3d();
  std::vector<float> vm = read_f32(arg("meas"), bins3d(m)), ve = read_f32(arg("est"), bins3d(e));
  sino3d_from(m, vm.data());
  sino3d_from(e, ve.data());
  Projector* pp = make_projector();
  {
  GuardedPixels guarded(&img);
  pp->DivideAndBackproject(m, e, &img);
  }
  write_f32(arg("out"), img.getPixelsPtr(), (size_t)img.getPixelCount());
  return 0;
}

// Optional sinogram argument `key` (raw f32 in the layout of `like`): a sinogram
// of the same class and geometry holding it, or NULL.
Sinogram3D* optional_sino(const char* key, Sinogram3D* like) {
  if (!has_arg(key)) return nullptr;
  Sinogram3D* c = like->Copy();
  std::vector<float> v = read_f32(arg(key), bins3d(like));
  sino3d_from(c, v.data());
  return c;
}

int cmd_osem3d() {
  Image init(size3d());
  size_t npix = (size_t)init.getPixelCount();
  std::vector<float> x0 = read_f32(arg("init"), npix);
  memcpy(init.getPixelsPtr(), x0.data(), npix * 4);
  Sinogram3DCylindricalPet* s = make_sino3d();
  std::vector<float> y = read_f32(arg("meas"), bins3d(s));
  sino3d_from(s, y.data());
  // guard=1: the reference's projector with its images in zeroed guard zones
  Projector* pp = (has_arg("guard") && iarg("guard")) ? (Projector*)new GuardedSiddon()
                                                      : make_projector();
  std::string dir = arg("outdir");
  int iters = iarg("iters"), subsets = iarg("subsets");
  OsemProbe osem(s, &init, dir, dir + "/osem3d", iters, 0, false, false, pp, pp, subsets);
  // add= / mult=: additive and multiplicative sinograms (full layout), each held in a
  // cylindrical sinogram built like the input (make_sino3d keeps the scanner radius)
  Sinogram3D* add = nullptr;
  Sinogram3D* mult = nullptr;
  if (has_arg("add")) {
    add = make_sino3d();
    std::vector<float> v = read_f32(arg("add"), bins3d(add));
    sino3d_from(add, v.data());
  }
  if (has_arg("mult")) {
    mult = make_sino3d();
    std::vector<float> v = read_f32(arg("mult"), bins3d(mult));
    sino3d_from(mult, v.data());
  }
  osem.corrections(add, mult);
  double t0 = now_s();
  if (!osem.Reconstruct()) return 6;
  report_time("osem3d", t0);
  write_f32((dir + "/final.f32").c_str(), osem.getReconstructedImage()->getPixelsPtr(), npix);
  std::vector<float> sens(npix * subsets);
  for (int k = 0; k < subsets; k++) memcpy(sens.data() + k * npix, osem.sens(k)->getPixelsPtr(), npix * 4);
  write_f32((dir + "/sens.f32").c_str(), sens.data(), sens.size());
  write_f32((dir + "/lik.f32").c_str(), osem.lik(), iters * subsets + 1);
  return 0;
}

int cmd_mlem3d() {
  Image init(size3d());
  size_t npix = (size_t)init.getPixelCount();
  std::vector<float> x0 = read_f32(arg("init"), npix);
  memcpy(init.getPixelsPtr(), x0.data(), npix * 4);
  Sinogram3DCylindricalPet* s = make_sino3d();
  std::vector<float> y = read_f32(arg("meas"), bins3d(s));
  sino3d_from(s, y.data());
  Projector* pp = make_projector();
  std::string dir = arg("outdir");
  int iters = iarg("iters");
  Mlem3dProbe mlem(s, &init, dir, dir + "/mlem3d", iters, 0, false, false, pp, pp);
  double t0 = now_s();
  if (!mlem.Reconstruct()) return 6;
  report_time("mlem3d", t0);
  write_f32((dir + "/final.f32").c_str(), mlem.getReconstructedImage()->getPixelsPtr(), npix);
  write_f32((dir + "/sens.f32").c_str(), mlem.sens()->getPixelsPtr(), npix);
  write_f32((dir + "/lik.f32").c_str(), mlem.lik(), iters + 1);
  return 0;
}


// Reference Siddon on a sample of rays of a 3D geometry: bins b = start + k*stride.
// For each sampled bin writes (bin, list length, sum of weights, projection of img
// accumulated with SiddonProjector::Project's rule, SiddonProjector.cpp:413-423).
int cmd_rays3d() {
  Image img(size3d());
  size_t npix = (size_t)img.getPixelCount();
  std::vector<float> x = read_f32(arg("img"), npix);
  memcpy(img.getPixelsPtr(), x.data(), npix * 4);
  SizeImage sz = img.getSize();
  Sinogram3DCylindricalPet* s = make_sino3d();
  float w = s->getAxialFoV_mm() / s->getNumRings();
  long start = atol(arg("start")), stride = atol(arg("stride"));
  std::vector<float> out;
  long bin = 0, next = start;
  for (int i = 0; i < s->getNumSegments(); i++)
    for (int j = 0; j < s->getSegment(i)->getNumSinograms(); j++) {
      Sinogram2DinCylindrical3Dpet* s2 = s->getSegment(i)->getSinogram2D(j);
      float z1 = (s2->getAxialValue1FromList(0) + s2->getAxialValue1FromList(s2->getNumZ() - 1)) / 2;
      float z2 = (s2->getAxialValue2FromList(0) + s2->getAxialValue2FromList(s2->getNumZ() - 1)) / 2;
      for (int k = 0; k < s2->getNumProj(); k++)
        for (int l = 0; l < s2->getNumR(); l++, bin++) {
          if (bin != next) continue;
          next += stride;
          Point3D p1, p2;
          float geom;
          s2->getPointsFromLor(k, l, 0, &p1, &p2, &geom);
          p1.Z = z1 - w / 2 + w / 2 + 0 * w;
          p2.Z = z2 - w / 2 + w / 2 + 0 * w;
          Line3D lor;
          lor.P0 = p1;
          lor.Vx = p2.X - p1.X;
          lor.Vy = p2.Y - p1.Y;
          lor.Vz = p2.Z - p1.Z;
          SiddonSegment* wl = nullptr;
          int len = 0;
          Siddon(lor, &img, &wl, &len, 1);
          float sw = 0, acc = 0;
          for (int n = 0; n < len; n++) {
            sw += wl[n].Segment;
            if (wl[n].IndexX < sz.nPixelsX && wl[n].IndexY < sz.nPixelsY && wl[n].IndexZ < sz.nPixelsZ) {
              long idx = (long)wl[n].IndexZ * sz.nPixelsX * sz.nPixelsY + (long)wl[n].IndexY * sz.nPixelsX + wl[n].IndexX;
              float pix = (idx >= 0 && idx < (long)npix) ? x[idx] : 0.0f;
              acc += wl[n].Segment * geom * pix;
            }
          }
          if (len > 0) free(wl);
          float rec[4];
          memcpy(&rec[0], &bin, 4);  // low 32 bits of the bin index
          rec[1] = (float)len;
          rec[2] = sw;
          rec[3] = acc;
          out.insert(out.end(), rec, rec + 4);
        }
    }
  write_f32(arg("out"), out.data(), out.size());
  return 0;
}


// One OSEM sub-iteration of the reference (OsemSinogram3d.cpp:164-237 minus the
// per-iteration file output): getSubset of the measured sinogram, Project,
// likelihood, inverseDivideBinToBin, Backproject, pixel update.  Timed as the CPU
// baseline of the OSEM workload.
int cmd_osem3d_sub() {
  Image img(size3d());
  size_t npix = (size_t)img.getPixelCount();
  for (size_t i = 0; i < npix; i++) img.getPixelsPtr()[i] = 1.0f;
  Sinogram3DCylindricalPet* s = make_sino3d();
  std::vector<float> y = read_f32(arg("meas"), bins3d(s));
  sino3d_from(s, y.data());
  int subsets = iarg("subsets"), sub = iarg("subset");
  Projector* pp = make_projector();
  Image bp(size3d());
  Image sens(size3d());
  {
    Sinogram3D* ones = s->getSubset(sub, subsets);
    ones->FillConstant(1);
    GuardedPixels gs(&sens);
    pp->Backproject(ones, &sens);
    delete ones;
  }
  double t0;
  float lik;
  {
    GuardedPixels gi(&img);
    GuardedPixels gb(&bp);
    t0 = now_s();
    Sinogram3D* in = s->getSubset(sub, subsets);
    Sinogram3D* est = s->getSubset(sub, subsets);
    est->FillConstant(0);
    pp->Project(&img, est);
    lik = est->getLikelihoodValue(in);
    bp.fillConstant(0);
    est->inverseDivideBinToBin(in);
    pp->Backproject(est, &bp);
    float* x = img.getPixelsPtr();
    float* b = bp.getPixelsPtr();
    float* sv = sens.getPixelsPtr();
    for (size_t k = 0; k < npix; k++) x[k] = (sv[k] >= 0) ? x[k] * b[k] / sv[k] : 0;
    report_time("osem3d_sub", t0);
    delete in;
    delete est;
  }
  printf("{\"likelihood\": %.9g}\n", lik);
  return 0;
}

}  // namespace


// ---- on-disk formats ---------------------------------------------------------
// img_write: an Image of the given size with pixels from `in`, written by the
// reference's Image::writeInterfile (medcon) to prefix.h33 / prefix.i33.
int cmd_img_write() {
  SizeImage sz = size3d();
  Image img(sz);
  std::vector<float> v = read_f32(arg("in"), (size_t)sz.nPixelsX * sz.nPixelsY * sz.nPixelsZ);
  memcpy(img.getPixelsPtr(), v.data(), v.size() * 4);
  if (!img.writeInterfile((char*)arg("prefix"))) {
    fprintf(stderr, "ref_driver: %s\n", img.getError().c_str());
    return 6;
  }
  return 0;
}

// img_read: Image::readFromInterfile of `header`; prints the size, writes the pixels.
int cmd_img_read() {
  Image img;
  if (!img.readFromInterfile((char*)arg("header"))) {
    fprintf(stderr, "ref_driver: %s\n", img.getError().c_str());
    return 6;
  }
  SizeImage sz = img.getSize();
  printf("{\"nx\": %d, \"ny\": %d, \"nz\": %d, \"dx\": %.9g, \"dy\": %.9g, \"dz\": %.9g}\n",
         sz.nPixelsX, sz.nPixelsY, sz.nPixelsZ, sz.sizePixelX_mm, sz.sizePixelY_mm,
         sz.sizePixelZ_mm);
  write_f32(arg("out"), img.getPixelsPtr(), (size_t)img.getPixelCount());
  return 0;
}

// sino3d_write: a cylindrical 3D sinogram (geometry as make_sino3d) with bins from
// `in`, written by Sinogram3D::writeInterfile to `header`.
int cmd_sino3d_write() {
  Sinogram3DCylindricalPet* s = make_sino3d();
  std::vector<float> v = read_f32(arg("in"), bins3d(s));
  sino3d_from(s, v.data());
  if (!s->writeInterfile(arg("header"))) return 6;
  return 0;
}

// sino3d_read: the reference's Sinogram3DCylindricalPet(interfile, rfov, zfov, rscanner)
// constructor; prints the segment table, writes the bins (segment-major).
int cmd_sino3d_read() {
  void* mem = calloc(1, sizeof(Sinogram3DCylindricalPet));
  Sinogram3DCylindricalPet* s = new (mem) Sinogram3DCylindricalPet(
      (char*)arg("header"), farg("rfov"), farg("zfov"), farg("rscanner"));
  printf("{\"nproj\": %d, \"nr\": %d, \"rings\": %d, \"nseg\": %d, \"nsino\": [", s->getNumProj(),
         s->getNumR(), s->getNumRings(), s->getNumSegments());
  for (int i = 0; i < s->getNumSegments(); i++)
    printf("%s%d", i ? ", " : "", s->getSegment(i)->getNumSinograms());
  printf("], \"minrd\": [");
  for (int i = 0; i < s->getNumSegments(); i++)
    printf("%s%d", i ? ", " : "", s->getSegment(i)->getMinRingDiff());
  printf("], \"maxrd\": [");
  for (int i = 0; i < s->getNumSegments(); i++)
    printf("%s%d", i ? ", " : "", s->getSegment(i)->getMaxRingDiff());
  printf("]}\n");
  std::vector<float> v = sino3d_to(s);
  write_f32(arg("out"), v.data(), v.size());
  return 0;
}

// par_read: parametersFile_read(file, method, key); prints the return code and value.
int cmd_par_read() {
  char value[512] = {0}, err[512] = {0};
  int rc = parametersFile_read((char*)arg("file"), (char*)arg("method"), (char*)arg("key"), value,
                               err);
  printf("%d\t%s\n", rc, value);
  return 0;
}

// itf_read: interfile_read(header, key); prints the return code and value.
int cmd_itf_read() {
  char value[1024] = {0}, err[1024] = {0};
  int rc = interfile_read((char*)arg("header"), (char*)arg("key"), value, err);
  printf("%d\t%s\n", rc, value);
  return 0;
}

// sino2d_write: a 2D sinogram (nproj x nr) with bins from `in`, written by
// Sinogram2D::writeInterfile to prefix.h33 / prefix.i33.
int cmd_sino2d_write() {
  Sinogram2DinCylindrical3Dpet sino(iarg("nproj"), iarg("nr"), farg("rfov"), farg("rscanner"));
  std::vector<float> v = read_f32(arg("in"), (size_t)iarg("nproj") * iarg("nr"));
  memcpy(sino.getSinogramPtr(), v.data(), v.size() * 4);
  if (!sino.writeInterfile(arg("prefix"))) return 6;
  return 0;
}

// sino2d_read: Sinogram2DinCylindrical3Dpet(header, rfov, rscanner); prints the size,
// writes the bins.
int cmd_sino2d_read() {
  Sinogram2DinCylindrical3Dpet sino((char*)arg("header"), farg("rfov"), farg("rscanner"));
  printf("{\"nproj\": %d, \"nr\": %d}\n", sino.getNumProj(), sino.getNumR());
  write_f32(arg("out"), sino.getSinogramPtr(), (size_t)sino.getNumProj() * sino.getNumR());
  return 0;
}

// mmr3d: a Siemens mMR sinogram read from `header` (Sinogram3DSiemensMmr(char*): mMR
// segments and LOR geometry, its bins as read); the image nx ny nz dx dz from `img` is
// projected into out_proj and the sinogram's bins are backprojected into out_back.
int cmd_mmr3d() {
  void* mem = calloc(1, sizeof(Sinogram3DSiemensMmr));
  Sinogram3DSiemensMmr* s = new (mem) Sinogram3DSiemensMmr((char*)arg("header"));
  Image img(size3d());
  size_t npix = (size_t)img.getPixelCount();
  std::vector<float> x = read_f32(arg("img"), npix);
  memcpy(img.getPixelsPtr(), x.data(), npix * 4);
  Image back(size3d());
  Projector* pp = make_projector();
  {
    GuardedPixels guarded(&back);
    pp->Backproject(s, &back);
  }
  write_f32(arg("out_back"), back.getPixelsPtr(), npix);
  {
    GuardedPixels guarded(&img);
    pp->Project(&img, s);
  }
  std::vector<float> v = sino3d_to(s);
  write_f32(arg("out_proj"), v.data(), v.size());
  return 0;
}

// mmr_mlem3d: MlemSinogram3d on a Siemens mMR sinogram read from `header`.  Its
// Copy() builds mMR segments with the scanner's constants, so the driver's copies
// keep their geometry (the cylindrical copy constructor does not), and the
// reconstruction is well defined.  The projector is GuardedSiddon (or, with
// projector=..., make_projector's): without the guard, LORs through the row before
// the image read heap memory, and the estimated bins of those LORs (and the
// likelihood) depend on what lies there.
int cmd_mmr_mlem3d() {
  void* mem = calloc(1, sizeof(Sinogram3DSiemensMmr));
  Sinogram3DSiemensMmr* s = new (mem) Sinogram3DSiemensMmr((char*)arg("header"));
  Image init(size3d());
  size_t npix = (size_t)init.getPixelCount();
  std::vector<float> x0 = read_f32(arg("init"), npix);
  memcpy(init.getPixelsPtr(), x0.data(), npix * 4);
  Projector* pp = has_arg("projector") ? make_projector() : (Projector*)new GuardedSiddon();
  std::string dir = arg("outdir");
  int iters = iarg("iters");
  Mlem3dProbe mlem(s, &init, dir, dir + "/mmr_mlem3d", iters, 0, false, false, pp, pp);
  // add= / mult=: additive and multiplicative sinograms; the mMR Copy() keeps the
  // scanner geometry (Sinogram3DSiemensMmr.cpp:74-87)
  mlem.corrections(optional_sino("add", s), optional_sino("mult", s));
  if (!mlem.Reconstruct()) return 6;
  write_f32((dir + "/final.f32").c_str(), mlem.getReconstructedImage()->getPixelsPtr(), npix);
  write_f32((dir + "/sens.f32").c_str(), mlem.sens()->getPixelsPtr(), npix);
  write_f32((dir + "/lik.f32").c_str(), mlem.lik(), iters + 1);
  return 0;
}

// mmr_osem3d: OsemSinogram3d on a Siemens mMR sinogram read from `header`, with
// `subsets` subsets.  Its subsets are Sinogram3DSiemensMmr::getSubset copies
// (Sinogram3DSiemensMmr.cpp:150-195), whose odd bins take half of the subset's own
// angular step (Sinogram2DinSiemensMmr.cpp:180-181).  add= / mult= as mmr_mlem3d.
int cmd_mmr_osem3d() {
  void* mem = calloc(1, sizeof(Sinogram3DSiemensMmr));
  Sinogram3DSiemensMmr* s = new (mem) Sinogram3DSiemensMmr((char*)arg("header"));
  Image init(size3d());
  size_t npix = (size_t)init.getPixelCount();
  std::vector<float> x0 = read_f32(arg("init"), npix);
  memcpy(init.getPixelsPtr(), x0.data(), npix * 4);
  Projector* pp = has_arg("projector") ? make_projector() : (Projector*)new GuardedSiddon();
  std::string dir = arg("outdir");
  int iters = iarg("iters"), subsets = iarg("subsets");
  OsemProbe osem(s, &init, dir, dir + "/mmr_osem3d", iters, 0, false, false, pp, pp, subsets);
  osem.corrections(optional_sino("add", s), optional_sino("mult", s));
  if (!osem.Reconstruct()) return 6;
  write_f32((dir + "/final.f32").c_str(), osem.getReconstructedImage()->getPixelsPtr(), npix);
  std::vector<float> sens(npix * subsets);
  for (int k = 0; k < subsets; k++) memcpy(sens.data() + k * npix, osem.sens(k)->getPixelsPtr(), npix * 4);
  write_f32((dir + "/sens.f32").c_str(), sens.data(), sens.size());
  write_f32((dir + "/lik.f32").c_str(), osem.lik(), iters * subsets + 1);
  return 0;
}

// mmr_subset_proj: Project of `img` onto subset `subset` of `subsets` of the mMR
// sinogram read from `header` (Sinogram3DSiemensMmr::getSubset), as the reference's
// OSEM projects it.
int cmd_mmr_subset_proj() {
  void* mem = calloc(1, sizeof(Sinogram3DSiemensMmr));
  Sinogram3DSiemensMmr* s = new (mem) Sinogram3DSiemensMmr((char*)arg("header"));
  Image img(size3d());
  size_t npix = (size_t)img.getPixelCount();
  std::vector<float> x = read_f32(arg("img"), npix);
  memcpy(img.getPixelsPtr(), x.data(), npix * 4);
  Sinogram3D* sub = s->getSubset(iarg("subset"), iarg("subsets"));
  Projector* pp = make_projector();
  {
    GuardedPixels guarded(&img);
    pp->Project(&img, sub);
  }
  std::vector<float> v = sino3d_to(sub);
  write_f32(arg("out"), v.data(), v.size());
  return 0;
}

// log_row: Logger(out).writeRowOfNumbers of the n floats in `in`.
int cmd_log_row() {
  const int n = iarg("n");
  std::vector<float> v = read_f32(arg("in"), (size_t)n);
  Logger logger(arg("out"));
  logger.writeRowOfNumbers(v.data(), n);
  return 0;
}

int main(int argc, char** argv) {
  // Keep every allocation on the brk heap: the reference's images are allocated
  // inside its EM drivers, where the guard of GuardedPixels cannot reach, and an
  // entry index of -1 at slice 0 reads one row before the array
  // (SiddonProjector.cpp:417).  With a page-aligned mmap'd array that read faults;
  // on the heap it reads a neighbouring chunk, and its weight is exactly 0
  // (the first crossing equals alpha_min there).
  mallopt(M_MMAP_THRESHOLD, 1 << 30);
  if (argc < 2) {
    fprintf(stderr, "usage: ref_driver <command> key=value ...\n");
    return 1;
  }
  for (int i = 2; i < argc; i++) {
    const char* eq = strchr(argv[i], '=');
    if (!eq) continue;
    g_args[std::string(argv[i], eq - argv[i])] = eq + 1;
  }
  std::string c = argv[1];
  if (c == "mmr_osem3d") return cmd_mmr_osem3d();
  if (c == "mmr_subset_proj") return cmd_mmr_subset_proj();
  if (c == "siddon2d") return cmd_siddon2d();
  if (c == "siddon3d") return cmd_siddon3d();
  if (c == "proj2d") return cmd_proj2d();
  if (c == "back2d") return cmd_back2d();
  if (c == "mlem2d") return cmd_mlem2d();
  if (c == "divback2d") return cmd_divback2d();
  if (c == "mlem2dmulti") return cmd_mlem2dmulti();
  if (c == "divback3d") return cmd_divback3d();
  if (c == "proj3d") return cmd_proj3d();
  if (c == "back3d") return cmd_back3d();
  if (c == "osem3d") return cmd_osem3d();
  if (c == "mlem3d") return cmd_mlem3d();
  if (c == "mmr_mlem3d") return cmd_mmr_mlem3d();
  if (c == "rays3d") return cmd_rays3d();
  if (c == "osem3d_sub") return cmd_osem3d_sub();
  if (c == "img_write") return cmd_img_write();
  if (c == "img_read") return cmd_img_read();
  if (c == "sino3d_write") return cmd_sino3d_write();
  if (c == "sino3d_read") return cmd_sino3d_read();
  if (c == "par_read") return cmd_par_read();
  if (c == "itf_read") return cmd_itf_read();
  if (c == "log_row") return cmd_log_row();
  if (c == "mmr3d") return cmd_mmr3d();
  if (c == "sino2d_write") return cmd_sino2d_write();
  if (c == "sino2d_read") return cmd_sino2d_read();
  fprintf(stderr, "ref_driver: unknown command %s\n", c.c_str());
  return 1;
}
