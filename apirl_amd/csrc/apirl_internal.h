// Internal types shared by the host side (capi.cpp, geometry.cpp) and the gfx950
// kernels (siddon_kernels.hip).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "apirl_hip.h"

namespace apirl {

// Per-LOR (angle, r) part of the Siddon setup that depends only on the transaxial
// geometry: recon/src/Siddon.cpp:83-145 (square-FOV alpha clip and clamp),
// :148-152/:494-498 (entry/exit points), :177-181/:509-512 (entry indices),
// :253-286/:583-606 (alpha increments and first crossings) and the xy part of the
// ray length (:248-250/:579-580).  Computed on the host in float with the reference's
// operation order and uploaded once; 64 bytes so one lane reads one line.
struct alignas(16) RayXY {
  float p0x, p0y;  // LOR.P0 (P1 of the bin)
  float vx, vy;    // LOR.Vx, LOR.Vy = P2 - P1
  float amin, amax;
  float ax, ay;    // first x / y crossing (2D: negative values clamped to amin)
  float axu, ayu;  // |dx / Vx|, |dy / Vy|
  float lxy;       // ((P0.X+Vx)-P0.X)^2 + ((P0.Y+Vy)-P0.Y)^2
  int i0, j0;      // entry voxel indices (may be -1 or n)
  int nxy;         // |imax - imin| + |jmax - jmin|
  int flags;       // see RAY_* below
  int pad;
};
static_assert(sizeof(RayXY) == 64, "RayXY must be one 64-byte line");

enum : int {
  RAY_VALID = 1,     // 2D: r <= rfov (Sinogram2DinCylindrical3Dpet.cpp:256-260)
  RAY_REJECTED = 2,  // 2D: both ends outside on one side (Siddon.cpp:515-519)
  RAY_XMAJOR = 64,   // |Vx| > |Vy|: neighbouring radial bins are offset along y
};
__host__ __device__ inline int ray_iinc(int flags) { return ((flags >> 2) & 3) - 1; }
__host__ __device__ inline int ray_jinc(int flags) { return ((flags >> 4) & 3) - 1; }
inline int ray_flags(int iinc, int jinc, int valid, int rejected) {
  return (valid ? RAY_VALID : 0) | (rejected ? RAY_REJECTED : 0) | ((iinc + 1) << 2) |
         ((jinc + 1) << 4);
}

// Geometry constants passed by value to every kernel.
struct Geo {
  int nx, ny, nz;
  float dx, dy, dz;
  float R;     // (dx * nx) / 2, Images.cpp:55
  float offz;  // dz / 2, Siddon.cpp:162
  int nr, nang, nsino;
  int dims;
  float r0, dr, rmargin;  // uniform radial grid r_j = r0 + j dr (gather candidate search)
  int walk24;  // voxel indices fit 24-bit multiplies (nx ny, nx, ny, nz below 2^23)
  int ns, nas;  // LORs per bin: transaxial x axial samples (oversampled 3D Siddon)
};

// A run of consecutive sinograms with increasing mean axial coordinate (one segment
// of the reference's enumeration): zmean(first + l) ~ z0 + l h.  The gather
// backprojector starts its sinogram scan at the index this line predicts for a ray
// of typical axial slope dlo (the middle sinogram's zb - za); dhi holds 1 / h.  The
// tile backprojector bounds the run's axial slopes zb - za by [dmin, dmax] and the
// deviation of zmean from the line by devz.
struct GatherSeg {
  int first, count;
  float z0, h, dlo, dhi, margin;
  int pad;
  float dmin, dmax, devz;
  int pad2;
};

// One geometry of a plan's LOR endpoints (APIRL_SCANNER_*) with the tables the
// projectors and the gather read for it.  A plan keeps every form it has built and
// points its active tables (apirl_plan::d_rays ...) at one of them.
struct LorForm {
  RayXY* rays = nullptr;        // [nang][nr][ns]
  float* xytab = nullptr;       // [nang][nr][nx + ny] gather entry crossings, or null
  float2* angcs_odd = nullptr;  // [nang] odd bins' (cos, sin) (radial lookup search)
  int2* rlut = nullptr;         // radial position -> bin range; null on a uniform grid
  float rl0 = 0.f, rl_inv_h = 0.f;
  int nrl = 0;
  bool built = false;
  bool gather = false;          // the gather's tables exist for this form
  // APIRL_PREC_F64_SETUP tables of the gather's column kernel, built on its first
  // use: the rays with the double setup's entry voxel, list length, alpha range and
  // first crossings (rounded to float), and their entry-crossing table
  RayXY* rays_f64s = nullptr;
  float* xytab_f64s = nullptr;
};

}  // namespace apirl

struct apirl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // this device's copy of the file-scope partials array of apirl_project_ratio_noalloc
  // (hipGetSymbolAddress, taken once per device); null if that failed
  double* d_noalloc = nullptr;
  // scratch owned by the context (likelihood partials, reductions)
  double* d_red = nullptr;
  size_t red_cap = 0;
  float* d_fred = nullptr;
  unsigned long long* d_ured = nullptr;
};

struct apirl_plan {
  apirl_ctx* ctx = nullptr;
  apirl_image_desc img{};
  apirl_sino_desc sino{};   // pointers inside are not kept
  apirl::Geo geo{};
  apirl::RayXY* d_rays = nullptr;  // [nang][nr][ns]
  float2* d_zab = nullptr;         // [nsino][nas] (P1.Z, P2.Z), 3D only
  // [16][nvox] per-XCD scatter targets, 8 in the image layout and 8 transposed
  // ([z][x][y]), allocated on the first scatter backprojection and zero between calls
  float* d_replicas = nullptr;
  bool replicas_failed = false;    // their allocation failed: direct atomics on the image
  // Oversampled plans (ns x nas LORs per bin): one plan of one LOR per bin for each
  // radial sub-LOR o and axial sample q (index o * nas + q), whose gathers add up to
  // the oversampled backprojection; built on the first gather, freed with the plan or
  // when the oversampling changes
  std::vector<apirl_plan*> os_parts;
  float* d_imgT = nullptr;         // [nz][nx][ny] transposed copy of the projected image
  int* d_zorder = nullptr;         // [nsino] sinograms sorted by mean axial coordinate
  float2* d_angcs = nullptr;       // [nang] (cos, sin) for the gather candidate search
  apirl::GatherSeg* d_segs = nullptr;
  int nseg = 0;
  bool gather_ok = false;          // uniform radial grid and increasing runs found
  int gather_zlanes = 1;           // 3D gather voxel mapping (1: 16x16 slice tiles; Z>1:
                                   // Z slices x 64/Z columns per wavefront)
  int gather_zvox = 3;             // 3D gather voxels per lane along z (1..4, tiles only)
  int gather_xcd = APIRL_GATHER_XCD_WEDGES;  // column kernel's tile order
  int tile_zs = 0;                 // tile backprojector's slab depth (0: about 5 slices)
  int* d_tile_order = nullptr;     // its item order (built for tile_order_n items)
  int tile_order_n = 0;
  // its geometry per (LOR form, subset, subsets), kept from the first call on, within
  // tile_geo_budget bytes in all (apirl_tuning.tile_cache_mb); `ready` once the launches
  // that build it are enqueued on `stream`, whose event `ev` follows them; `list` the
  // sinogram lists of its (item, ray) pairs (lists: 0 not tried, 1 kept, -1 beyond the
  // budget)
  struct TileGeo {
    int form, subset, nsubsets;
    void* d;
    size_t bytes;
    bool ready;
    hipStream_t stream;
    hipEvent_t ev;
    unsigned short* list;
    size_t list_bytes;
    int lists;
  };
  std::vector<TileGeo> tile_geo;
  size_t tile_geo_used = 0;
  size_t tile_geo_budget = (size_t)16384 << 20;
  // its tile maps (APIRL_GATHER_XCD_WEDGES), one per voxels-per-lane depth 1..4, built
  // on first use and kept until the plan ends (captured launches may hold them)
  int* d_tilemap[5] = {};
  int tilemap_slots[5] = {};
  void* d_gscratch = nullptr;      // per-ray axial setup of the gather (grown on demand)
  float* d_divbins = nullptr;      // DivideAndBackproject's bin values for the gather
  size_t divbins_cap = 0;          //   (grown on demand)
  int* d_zsoff = nullptr;          // [2 nsino + 1] z-crossing table slots, first slices
  float* d_xytab = nullptr;        // [nang][nr][nx + ny] entry crossings (gather)
  int zblock = 0;                  // z-crossing table entries per (angle, radial bin)
  size_t gscratch_bytes = 0;
  std::vector<float> ang, r;       // host copies of the grids
  std::vector<float2> zab;         // [nsino] one axial sample per bin
  std::vector<float> z1m, z2m;     // [nsino] mean ring coordinates of each sinogram
  bool gather_tables = false;      // the gather's candidate-search tables common to every
                                   // LOR form (angle directions, segment runs, z slots)
  bool uniform_r = true;           // the plan's radial grid is uniform (closed-form search)
  float2* d_angcs_odd = nullptr;   // [nang] odd bins' (cos, sin) (mMR gather search)
  int2* d_rlut = nullptr;          // [nrl] radial position -> bin range (mMR gather search)
  float rl0 = 0.f, rl_inv_h = 0.f;
  int nrl = 0;
  int scanner = 0;                 // APIRL_SCANNER_*: geometry of the LOR endpoints
  apirl::LorForm forms[3];         // per APIRL_SCANNER_*, built on first use; the d_rays,
                                   // d_xytab, d_angcs_odd, d_rlut above alias forms[scanner]
};

namespace apirl {
// Points the plan's active LOR tables at form f (built); returns the previous form.
int use_lor_form(apirl_plan* p, int f);
// Builds form f of a 3D plan if needed (host tables, upload).  false: out of memory or
// not defined for the plan (message in the context).
bool build_lor_form(apirl_plan* p, int f);
// Holds form f active for a scope (host-side pointer swap: launches already enqueued
// keep the tables they were given).
struct ActiveForm {
  apirl_plan* p;
  int prev;
  ActiveForm(apirl_plan* plan, int f) : p(plan), prev(use_lor_form(plan, f)) {}
  ~ActiveForm() { use_lor_form(p, prev); }
};
}  // namespace apirl

namespace apirl {

// geometry.cpp
void build_ray_table(const apirl_image_desc& img, const apirl_sino_desc& sino,
                     const std::vector<float>& ang, const std::vector<float>& r,
                     std::vector<RayXY>& out);
void build_z_table(const apirl_sino_desc& sino, std::vector<float2>& zab);
// Ray records from explicit transaxial LOR endpoints (x1, y1, x2, y2) per bin; ok[n]
// false marks a LOR the geometry rejects.
void build_ray_table_pts(const apirl_image_desc& img, bool is2d, const std::vector<float4>& pts,
                         const std::vector<char>& ok, std::vector<RayXY>& out);
// Siemens Biograph mMR LOR endpoints of a nang x nr sinogram.
void mmr_lor_points(int nang, int nr, std::vector<float4>& pts);
// The LORs of the reference's mMR subset sinograms (APIRL_SCANNER_MMR_SUBSET):
// cylindrical, on the radius 328 mm, at the mMR's radial positions and angles.
void mmr_subset_lor_points(int nang, int nr, std::vector<float4>& pts);
// Oversampled bins (SiddonProjector.cpp:374-395, Sinogram2DinCylindrical3Dpet.cpp:
// 283-299): ns radial sub-LORs per bin and nas axial ones per sinogram.
void build_subsample_r(const std::vector<float>& r, int ns, std::vector<float>& r_sub);
void build_z_table_os(const std::vector<float>& z1m, const std::vector<float>& z2m,
                      float ring_width, int nas, std::vector<float2>& zab);
int subset_num_angles(int nang, int nsubsets, int subset);

// siddon_kernels.hip: launchers (enqueue on stream; return hipError_t)
struct FusedRatio {
  const float* meas = nullptr;      // full sinogram
  const float* additive = nullptr;  // full sinogram or null
  float* ratio = nullptr;           // subset sinogram
  double* lik_partials = nullptr;   // one per block, or null
  int lik_positive_only = 0;
  // angle of the additive sinogram added to bin (s, k, r) of subset `subset`:
  // 0 = the bin's own angle subset + nsubsets k; 1 = k, the subset-local index into
  // the full additive sinogram, as OsemSinogram3d.cpp:172-173 with
  // Sinogram3D::addBinToBin (Sinogram3D.cpp:857-873) reads it
  int additive_local = 0;
};
// 3D, reference precision, one LOR per bin (path_forward_ok): the path forward
// projector, which writes project_partials(.., true) likelihood partials (one per
// subset (angle, bin)) instead of one per 256 bins.
hipError_t launch_project(const apirl_plan* p, const float* img, float* sino, int subset,
                          int nsubsets, int precision, const FusedRatio* fused, int* nblocks_out,
                          hipStream_t st);
bool path_forward_ok(const apirl_plan* p, int precision);
int project_partials(const apirl_plan* p, int subset, int nsubsets, bool path);
hipError_t launch_backproject(const apirl_plan* p, const float* sino, const float* est,
                              float* img, int subset, int nsubsets, int kind, int mode,
                              hipStream_t st);
// zeroes the scatter's n-float image copies (device-scope stores, as the reduction
// clears them)
hipError_t launch_zero_replicas(float* rep, size_t n, hipStream_t st);
hipError_t launch_em_update(float* x, const float* b, const float* s, size_t n, float thr,
                            int strict, hipStream_t st);
hipError_t launch_force_positive(float* x, size_t n, hipStream_t st);
// DivideAndBackproject's per-bin value (3D and 2D rules) into out, for its gather form.
hipError_t launch_divide_rule(const float* meas, const float* est, float* out, size_t n,
                              int dims, hipStream_t st);
hipError_t launch_sum_partials(const double* partials, int n, double* out, hipStream_t st);
// Partials of apirl_project_ratio_noalloc: a file-scope device array of
// siddon_kernels.hip of kNoallocPartials doubles (one per workgroup of 256 bins).
constexpr int kNoallocPartials = 1 << 19;
// The current device's address of that array (hipGetSymbolAddress).
hipError_t noalloc_scratch_address(double** out);
// The one-LOR-per-bin plans of an oversampled plan (apirl_plan::os_parts); false with
// the context's error set when they cannot be built.
bool ensure_os_parts(apirl_plan* p);
// The rays of an APIRL_PREC_F64_SETUP gather: each ray's transaxial setup carried out
// in double from its float LOR (setup_xy_f64), rounded to float as the projector's
// F64_SETUP loop starts from it.
void build_f64s_rays(const Geo& g, const std::vector<RayXY>& rays, std::vector<RayXY>& out);
void free_os_parts(apirl_plan* p);
// Likelihood sum y log(x) - x over x > 0 (positive_only) or x != 0, deterministic:
// block partials into `partials`, then a fixed-order sum into *out.
hipError_t launch_likelihood(const float* x, const float* y, long n, int positive_only,
                             double* partials, double* out, hipStream_t st);
hipError_t launch_min_nz_max(const float* x, size_t n, float* d_out2, hipStream_t st);
hipError_t launch_count_intersections(const apirl_plan* p, int subset, int nsubsets,
                                      unsigned long long* d_out, hipStream_t st);

// The EM update a gather can apply to its voxels instead of storing their
// backprojection (GatherArgs::upd_sens): img is then the image being updated.
struct EmFuse {
  const float* sens;
  float thr;
  int strict;
};
hipError_t launch_backproject_gather(const apirl_plan* p, const float* sino, float* img,
                                     int subset, int nsubsets, int kind, int precision,
                                     void* scratch, size_t scratch_bytes, hipStream_t st,
                                     const EmFuse* fuse = nullptr);
// Tile-walk backprojector (APIRL_BP_TILE); tile_misfit says why a plan cannot take it
// (null when it can).  Its geometry for one subset (the candidate rays of every tile
// and their paths; tile_geo_bytes) depends on the plan's LOR form alone: with `geo`
// null it is built in the scratch on every call; otherwise it lives in `geo`, built
// there when `geo_ready` is false and reused as it stands when true (with `list`, the
// sinogram lists launch_tile_lists kept for it, when not null).  The scratch
// (tile_scratch_bytes, with the geometry when `geo` is null) holds the call's axial
// records with their bin values and its live-ray lists.
hipError_t launch_backproject_tile(const apirl_plan* p, const float* sino, float* img,
                                   int subset, int nsubsets, int kind, void* scratch,
                                   size_t scratch_bytes, hipStream_t st, const EmFuse* fuse,
                                   void* geo, bool geo_ready,
                                   const unsigned short* list = nullptr);
// The sinogram lists of every (item, candidate ray) of a kept geometry (built):
// with list null, counts them (*total: the device address of their number, a long
// long, valid once the stream reaches it); otherwise writes them to list (2 bytes
// each).  launch_backproject_tile then reads them instead of listing.
hipError_t launch_tile_lists(const apirl_plan* p, int subset, int nsubsets, void* geo,
                             unsigned short* list, long long** total, hipStream_t st);
const char* tile_misfit(const apirl_plan* p);
size_t tile_geo_bytes(const apirl_plan* p, int subset, int nsubsets);
size_t tile_scratch_bytes(const apirl_plan* p, int subset, int nsubsets, bool with_geo);
int tile_items(const apirl_plan* p);
// item order of the tile backprojector: the tiles nearest the image centre first
void build_tile_order(const apirl_plan* p, std::vector<int>& order);
size_t gather_scratch_bytes(const apirl_plan* p, int subset, int nsubsets);
// The 3D gather's 32-bit record and table offsets hold for this subset (fewer than
// 2^31 subset bins; angle x radial bins, sinograms and crossing-table slots below 2^24).
bool gather_fits(const apirl_plan* p, int subset, int nsubsets);
// Tile order of the column kernel that keeps each XCD on one angular sector of the
// image (APIRL_GATHER_XCD_WEDGES): 8 lists of tiles, one per 45-degree wedge about
// the image centre, z-block by z-block, interleaved so that slot b goes to list
// b % 8, the XCD that the hardware dispatches workgroup b to; -1 pads the lists.
void build_wedge_tile_map(int ntx, int nty, int nzb, std::vector<int>& map);
// Why a subset does not fit the gather's 32-bit offsets (nullptr when it fits).
const char* gather_misfit(const apirl_plan* p, int subset, int nsubsets);
size_t ztable_bytes(const apirl_plan* p, int nang_sub);
void build_xy_crossings(const apirl_image_desc& img, const std::vector<RayXY>& rays,
                        std::vector<float>& out);
// Radial positions of the mMR's bins (ascending) and the (cos, sin) of each angle's
// odd bins (half an angular step less), as the mMR LOR endpoints use them.
void mmr_radial_grid(int nang, int nr, std::vector<float>& r, std::vector<float2>& angcs_odd);
// Lookup table from a radial position to the range of bins around it, on a grid of
// spacing h from rl0 covering the ascending positions r with two points to spare:
// lut[k] = (first j with r[j] >= rl0 + k h, last j with r[j] <= rl0 + k h).
void build_radial_lut(const std::vector<float>& r, std::vector<int2>& lut, float& rl0,
                      float& inv_h);
void build_gather_tables(const apirl_plan& p, std::vector<float2>& angcs,
                         std::vector<GatherSeg>& segs, float& r0, float& dr, float& rmargin,
                         bool& ok);

// Backprojection kinds
enum { BP_PLAIN = 0, BP_ONES = 1, BP_DIVIDE = 2, BP_MULT = 3 };

// ---- batched jobs (batch.cpp, grouped kernels) --------------------------------
// A round runs one OSEM / MLEM sub-iteration of every active job with one launch
// per phase; each launch walks a job table in device memory.  Every job owns a
// contiguous run of "virtual blocks" (first .. next job's first) of that phase;
// the launches are persistent (fixed grid) and their workgroups take virtual blocks
// from a per-phase counter (dynamic balance: tiles differ ~10x in cost), so the
// launch shapes never change and one captured graph serves every round.

// Per-ray axial part of the Siddon setup and the bin value of one subset bin, used
// by the gather backprojector (one 32-byte load per candidate ray).
struct alignas(32) ZRay {
  int k0;       // entry slice
  short k1;     // exit slice
  short kinc;   // sign of Vz
  float azu;    // |dz / Vz|
  float az0;    // first z crossing (FLT_MAX when Vz == 0)
  float L;      // ray length |P2 - P1| in the reference's operand order
  float b;      // bin value being backprojected
  int zcount;   // z crossings of the ray held by the crossing table (the rest: alpha_after)
  int pad;
};
static_assert(sizeof(ZRay) == 32, "ZRay is one 32-byte load");

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
  int zlanes;         // voxel mapping (see gather_kernel)
  // z-crossing table: for every subset ray, the reference's accumulated z crossing
  // by which it enters slice k, at ztab[((zsoff[s] + k - kmin[s]) * nang_sub + ka) * nr + r]
  float* ztab;
  const int* zsoff;   // [nsino + 1] slot offsets, then [nsino] first slice kmin per sinogram
  int zblock;
  // x / y crossing table of the plan: xytab[(a * nr + r) * (nx + ny) + i] is the
  // accumulated x crossing by which ray (a, r) enters column i, [... + nx + j] the y
  // crossing by which it enters row j
  const float* xytab;
  // Candidate search on a non-uniform radial grid with odd bins at their own angle
  // (mMR; null for the plan's uniform grid): angcs_odd[a] is the (cos, sin) of the
  // odd bins of angle a; rlut[k] = (first bin with r >= rl0 + k h, last bin with
  // r <= rl0 + k h), k < nrl, rl_inv_h = 1 / h.
  const float2* angcs_odd;
  const int2* rlut;
  float rl0, rl_inv_h;
  int nrl;
  // Column kernel's workgroup order (null: the 3D grid itself): slot b of a 1D grid
  // takes tile tile_map[b] = tx + ntx (ty + nty tz), or nothing when it is -1.
  const int* tile_map;
  int tile_slots;
  // EM update fused into the gather (null: img accumulates the backprojection): img is
  // then the image x, and each voxel's backprojection b, summed from zero, updates it
  // to (s > thr or, not strict, s >= thr) ? (x b) / s : 0 with s = upd_sens[v]
  const float* upd_sens;
  float upd_thr;
  int upd_strict;
};

struct BatchHeader {
  long total[8];                // virtual blocks of each phase (BatchPhase)
  int count[8];                 // jobs in each phase's table
  unsigned long long next[8];   // work counters (uploaded as 0 with every round's table)
  int grab;                     // virtual blocks a workgroup takes per counter draw (>= 1)
  int pad;
};
enum BatchPhase { PH_CLEAR = 0, PH_PROJ2, PH_PROJ3, PH_ZSETUP, PH_GATHER2, PH_GATHER3, PH_UPDATE };

struct ClearJob {  // zero n floats (or ints) at p
  float* p;
  long n;
  long first;
};
struct ProjJob {  // fused projection + ratio of one subset (likelihood not taken)
  Geo g;
  const RayXY* rays;
  const float2* zab;
  const int* zorder;
  const float* img;
  FusedRatio fr;
  int nang_sub, subset, nsubsets, pad;
  long nrays;
  long first;
};
struct ZsJob {  // axial records of one subset for the gather
  Geo g;
  const RayXY* rays;
  const float2* zab;
  const float* sino;
  ZRay* out;
  int* nz_pair;
  float* ztab;
  const int* zsoff;
  int zblock, pad;
  int kind, nang_sub, subset, nsubsets;
  long nrays;
  long first;
};
struct GJob {  // gather backprojection of one subset (16 x 16 slice tiles)
  GatherArgs A;
  int ntx, nty;
  long first;
};
struct UpdJob {  // x := (s cmp thr) ? x b / s : 0
  float* x;
  const float* b;
  const float* s;
  long n;
  float thr;
  int strict;
  long first;
};
constexpr int kGroupBlock = 256;  // threads per workgroup of the grouped launches
inline long group_blocks(long n) { return (n + kGroupBlock - 1) / kGroupBlock; }
// Voxels per lane along z of the grouped 3D gather (the column kernel's default).
constexpr int kGroupZV = 3;
inline long gather_tile_blocks(const Geo& g) {
  const int nzg = g.dims == 3 ? (g.nz + kGroupZV - 1) / kGroupZV : g.nz;
  return (long)((g.nx + 15) / 16) * ((g.ny + 15) / 16) * nzg;
}

hipError_t launch_group_clear(BatchHeader* hdr, const ClearJob* jobs, int grid,
                              hipStream_t st);
hipError_t launch_group_project(BatchHeader* hdr, const ProjJob* jobs2,
                                const ProjJob* jobs3, int grid, hipStream_t st);
hipError_t launch_group_update(BatchHeader* hdr, const UpdJob* jobs, int grid,
                               hipStream_t st);
hipError_t launch_group_gather(BatchHeader* hdr, const ZsJob* zs, const GJob* g2,
                               const GJob* g3, int grid, int max_nsino, hipStream_t st);
void gather_args_for(const apirl_plan* p, int subset, int nsubsets, int kind, const float* sino,
                     float* img, void* scratch, GatherArgs& A, ZRay*& zr, int*& nz_pair);
// LDS bytes of the 3D gather launches for nsino sinograms (axial table, slot offsets
// and first slices of the z-crossing table)
inline size_t gather_lds_bytes(int nsino) {
  return (size_t)nsino * sizeof(float2) + (size_t)(2 * nsino + 1) * sizeof(int);
}

}  // namespace apirl
