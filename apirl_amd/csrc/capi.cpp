// C-ABI of the HIP Siddon path: contexts, geometry plans, the Projector entry points
// and the device-resident MLEM / OSEM drivers.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "apirl_internal.h"

using apirl::Geo;

namespace {

apirl_status fail(apirl_ctx* ctx, apirl_status code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

apirl_status hip_fail(apirl_ctx* ctx, hipError_t e, const char* where) {
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  return fail(ctx, APIRL_ERR_HIP, m);
}

#define APIRL_HIP(ctx, call)                             \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

// Makes the plan's device current for the duration of a call.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

apirl_status ensure_red(apirl_ctx* ctx, size_t n) {
  if (ctx->red_cap >= n) return APIRL_OK;
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  ctx->d_red = nullptr;
  ctx->red_cap = 0;
  size_t cap = n < 1024 ? 1024 : n;
  APIRL_HIP(ctx, hipMalloc(&ctx->d_red, (cap + 8) * sizeof(double)));
  ctx->red_cap = cap;
  return APIRL_OK;
}

bool check_subset(const apirl_plan* p, int subset, int nsubsets) {
  return nsubsets >= 1 && subset >= 0 && subset < nsubsets && nsubsets <= p->geo.nang;
}

void free_f64s(apirl::LorForm& f) {
  if (f.rays_f64s) (void)hipFree(f.rays_f64s);
  if (f.xytab_f64s) (void)hipFree(f.xytab_f64s);
  f.rays_f64s = nullptr;
  f.xytab_f64s = nullptr;
}

void free_form(apirl::LorForm& f) {
  free_f64s(f);
  if (f.rays) (void)hipFree(f.rays);
  if (f.xytab) (void)hipFree(f.xytab);
  if (f.angcs_odd) (void)hipFree(f.angcs_odd);
  if (f.rlut) (void)hipFree(f.rlut);
  f = apirl::LorForm();
}

// The gather's tables of one LOR form: the entry-crossing table of its rays and,
// when the radial grid r is not the plan's uniform one, the radial lookup table and
// the odd bins' directions (half a step earlier when odd_half_step, the mMR's;
// the bin's own angle otherwise).  false leaves the form without gather tables.
bool upload_gather_form(apirl_plan* p, const std::vector<apirl::RayXY>& rays,
                        const std::vector<float>& r, bool lut, bool odd_half_step,
                        apirl::LorForm& f) {
  std::vector<float> xy;
  apirl::build_xy_crossings(p->img, rays, xy);
  bool ok = hipMalloc(&f.xytab, xy.size() * sizeof(float)) == hipSuccess &&
            hipMemcpy(f.xytab, xy.data(), xy.size() * sizeof(float), hipMemcpyHostToDevice) ==
                hipSuccess;
  if (ok && lut) {
    std::vector<float2> odd;
    std::vector<int2> tab;
    if (odd_half_step) {
      std::vector<float> rr;
      apirl::mmr_radial_grid(p->geo.nang, p->geo.nr, rr, odd);
    } else {
      odd.resize(p->geo.nang);
      for (int a = 0; a < p->geo.nang; a++) {
        const double rad = (double)p->ang[a] * 0.01745329251994329576;
        odd[a].x = (float)cos(rad);
        odd[a].y = (float)sin(rad);
      }
    }
    apirl::build_radial_lut(r, tab, f.rl0, f.rl_inv_h);
    f.nrl = (int)tab.size();
    ok = hipMalloc(&f.angcs_odd, odd.size() * sizeof(float2)) == hipSuccess &&
         hipMemcpy(f.angcs_odd, odd.data(), odd.size() * sizeof(float2),
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMalloc(&f.rlut, tab.size() * sizeof(int2)) == hipSuccess &&
         hipMemcpy(f.rlut, tab.data(), tab.size() * sizeof(int2), hipMemcpyHostToDevice) ==
             hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    apirl::RayXY* keep = f.rays;
    f.rays = nullptr;
    free_form(f);
    f.rays = keep;
  }
  return ok;
}

// The sector order of the gather column kernel's tiles (APIRL_GATHER_XCD_WEDGES) for
// the plan's voxels-per-lane depth: built once per depth and kept until the plan
// ends, so a launch captured with one map never sees it freed.  False when it cannot
// be allocated; 2D plans and the other orders need none.
bool ensure_tile_map(apirl_plan* p) {
  const int zv = p->gather_zvox;
  if (p->gather_xcd != APIRL_GATHER_XCD_WEDGES || p->geo.dims != 3 || zv < 1 || zv > 4 ||
      p->d_tilemap[zv])
    return true;
  std::vector<int> map;
  apirl::build_wedge_tile_map((p->geo.nx + 15) / 16, (p->geo.ny + 15) / 16,
                              (p->geo.nz + zv - 1) / zv, map);
  int* d = nullptr;
  if (hipMalloc(&d, map.size() * sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (hipMemcpy(d, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(d);
    return false;
  }
  p->d_tilemap[zv] = d;
  p->tilemap_slots[zv] = (int)map.size();
  return true;
}

// The APIRL_PREC_F64_SETUP tables of the active form's gather (LorForm::rays_f64s,
// xytab_f64s), built from its rays on first use: false when they cannot be, and the
// F64_SETUP gather then keeps its one-voxel kernel.
bool ensure_f64s_tables(apirl_plan* p) {
  apirl::LorForm& f = p->forms[p->scanner];
  if (f.rays_f64s && f.xytab_f64s) return true;
  if (!f.rays || p->geo.dims != 3 || p->geo.ns * p->geo.nas != 1) return false;
  std::vector<apirl::RayXY> rays((size_t)p->geo.nang * p->geo.nr);
  if (hipMemcpy(rays.data(), f.rays, rays.size() * sizeof(apirl::RayXY),
                hipMemcpyDeviceToHost) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  std::vector<apirl::RayXY> r64;
  apirl::build_f64s_rays(p->geo, rays, r64);
  std::vector<float> xy;
  apirl::build_xy_crossings(p->img, r64, xy);
  const bool ok =
      hipMalloc(&f.rays_f64s, r64.size() * sizeof(apirl::RayXY)) == hipSuccess &&
      hipMemcpy(f.rays_f64s, r64.data(), r64.size() * sizeof(apirl::RayXY),
                hipMemcpyHostToDevice) == hipSuccess &&
      hipMalloc(&f.xytab_f64s, xy.size() * sizeof(float)) == hipSuccess &&
      hipMemcpy(f.xytab_f64s, xy.data(), xy.size() * sizeof(float), hipMemcpyHostToDevice) ==
          hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    free_f64s(f);
  }
  return ok;
}

// The tile backprojector's item order for the plan's current slab depth, built on
// first use (and again when the slab depth changes the item count).
bool ensure_tile_order(apirl_plan* p) {
  const int n = apirl::tile_items(p);
  if (p->d_tile_order && p->tile_order_n == n) return true;
  std::vector<int> order;
  apirl::build_tile_order(p, order);
  if (p->d_tile_order) {
    if (hipStreamSynchronize(p->ctx->stream) != hipSuccess) return false;
    (void)hipFree(p->d_tile_order);
  }
  p->d_tile_order = nullptr;
  p->tile_order_n = 0;
  if (hipMalloc(&p->d_tile_order, order.size() * sizeof(int)) != hipSuccess ||
      hipMemcpy(p->d_tile_order, order.data(), order.size() * sizeof(int),
                hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    if (p->d_tile_order) (void)hipFree(p->d_tile_order);
    p->d_tile_order = nullptr;
    return false;
  }
  p->tile_order_n = n;
  return true;
}

// The kept tile geometry of (the plan's LOR form, subset, subsets): found, or allocated
// within the plan's budget (not ready until a call has built it).  Null when it does
// not fit the budget or the stream is being captured (a capture would only record the
// launches that build it): the call then builds it in its scratch.
apirl_plan::TileGeo* tile_geo_for(apirl_plan* p, int subset, int nsubsets) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(p->ctx->stream, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    return nullptr;
  }
  for (apirl_plan::TileGeo& e : p->tile_geo)
    if (e.form == p->scanner && e.subset == subset && e.nsubsets == nsubsets) return &e;
  const size_t bytes = apirl::tile_geo_bytes(p, subset, nsubsets);
  if (bytes == 0 || p->tile_geo_used + bytes > p->tile_geo_budget) return nullptr;
  apirl_plan::TileGeo e{p->scanner, subset, nsubsets, nullptr, bytes, false, nullptr, nullptr,
                        nullptr, 0, 0};
  if (hipMalloc(&e.d, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (hipEventCreateWithFlags(&e.ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(e.d);
    return nullptr;
  }
  p->tile_geo_used += bytes;
  p->tile_geo.push_back(e);
  return &p->tile_geo.back();
}

void free_tile_geo(apirl_plan* p) {
  for (apirl_plan::TileGeo& e : p->tile_geo) {
    (void)hipEventSynchronize(e.ev);
    (void)hipFree(e.d);
    if (e.list) (void)hipFree(e.list);
    (void)hipEventDestroy(e.ev);
  }
  p->tile_geo.clear();
  p->tile_geo_used = 0;
}

// Backprojection of one subset in either form.
// Oversampled bins are traversed in the reference's arithmetic only (and need the
// 24-bit index walk).
bool os_precision_ok(const apirl_plan* p, int precision) {
  return p->geo.ns * p->geo.nas == 1 || (precision == APIRL_PREC_REFERENCE && p->geo.walk24);
}

// fuse (gather form, one LOR per bin): img is the EM image, updated by each voxel's
// backprojection instead of accumulating it
apirl_status backproject_any(apirl_plan* p, const float* sino, const float* est, float* img,
                             int subset, int nsubsets, int kind, int mode,
                             int precision = APIRL_PREC_REFERENCE,
                             const apirl::EmFuse* fuse = nullptr) {
  apirl_ctx* c = p->ctx;
  if (precision < APIRL_PREC_REFERENCE || precision > APIRL_PREC_F64_SETUP)
    return fail(c, APIRL_ERR_INVALID, "invalid precision");
  if (mode == APIRL_BP_SCATTER) {
    if (precision != APIRL_PREC_REFERENCE)
      return fail(c, APIRL_ERR_UNSUPPORTED,
                  "the scatter backprojector has the reference precision only");
    if (!p->d_replicas && !p->replicas_failed) {
      // the per-XCD scatter targets (16 image copies), zeroed here and re-zeroed by
      // the reduction after every use; without them the atomics go to the image
      const size_t bytes = (size_t)16 * p->geo.nx * p->geo.ny * p->geo.nz * sizeof(float);
      if (hipMalloc(&p->d_replicas, bytes) != hipSuccess) {
        p->d_replicas = nullptr;
        p->replicas_failed = true;
        (void)hipGetLastError();
      } else if (apirl::launch_zero_replicas(p->d_replicas, bytes / sizeof(float), c->stream) !=
                 hipSuccess) {
        (void)hipFree(p->d_replicas);
        p->d_replicas = nullptr;
        return fail(c, APIRL_ERR_HIP, "cannot clear the scatter backprojector's copies");
      }
    }
    APIRL_HIP(c, apirl::launch_backproject(p, sino, est, img, subset, nsubsets, kind, mode,
                                           c->stream));
    return APIRL_OK;
  }
  if (mode != APIRL_BP_GATHER && mode != APIRL_BP_TILE)
    return fail(c, APIRL_ERR_INVALID, "invalid backprojection mode");
  if (fuse && p->geo.ns * p->geo.nas > 1)
    return fail(c, APIRL_ERR_INVALID,
                "the fused update needs the gather or tile form of one LOR per bin");
  if (mode == APIRL_BP_TILE) {
    if (precision != APIRL_PREC_REFERENCE)
      return fail(c, APIRL_ERR_UNSUPPORTED, "the tile backprojector has the reference precision only");
    if (const char* why = apirl::tile_misfit(p)) return fail(c, APIRL_ERR_UNSUPPORTED, why);
  }
  if (p->geo.ns * p->geo.nas > 1) {
    // every LOR of the oversampled bins (SiddonProjector.cpp:491-548): the gathers of
    // the plans of one LOR per bin, added in turn into the image
    if (precision != APIRL_PREC_REFERENCE)
      return fail(c, APIRL_ERR_UNSUPPORTED, "oversampled LORs use the reference precision");
    if (!apirl::ensure_os_parts(p)) return APIRL_ERR_UNSUPPORTED;
    for (apirl_plan* h : p->os_parts) {
      const apirl_status st = backproject_any(h, sino, est, img, subset, nsubsets, kind, mode,
                                              precision);
      if (st != APIRL_OK) return st;
    }
    return APIRL_OK;
  }
  if (!p->gather_ok)
    return fail(c, APIRL_ERR_UNSUPPORTED,
                "gather needs its tables (a radial grid it can search) and one LOR per bin "
                "(no oversampling)");
  if (const char* why = apirl::gather_misfit(p, subset, nsubsets))
    return fail(c, APIRL_ERR_UNSUPPORTED,
                std::string("gather: ") + why + " (use the scatter form)");
  apirl_plan::TileGeo* tg = mode == APIRL_BP_TILE ? tile_geo_for(p, subset, nsubsets) : nullptr;
  const size_t need = mode == APIRL_BP_TILE
                          ? apirl::tile_scratch_bytes(p, subset, nsubsets, tg == nullptr)
                          : apirl::gather_scratch_bytes(p, subset, nsubsets);
  if (need > p->gscratch_bytes) {
    if (p->d_gscratch) {
      APIRL_HIP(c, hipStreamSynchronize(c->stream));
      (void)hipFree(p->d_gscratch);
    }
    p->d_gscratch = nullptr;
    p->gscratch_bytes = 0;
    if (hipMalloc(&p->d_gscratch, need) != hipSuccess)
      return fail(c, APIRL_ERR_NOMEM, "cannot allocate the gather scratch");
    p->gscratch_bytes = need;
  }
  if (kind == apirl::BP_DIVIDE) {
    // DivideAndBackproject gathered: each bin's value by the reference's rule, then a
    // plain backprojection of it (the bins the rule zeroes add nothing, as the bins
    // the reference skips)
    const size_t nb = apirl_plan_subset_bins(p, subset, nsubsets);
    if (nb > p->divbins_cap) {
      if (p->d_divbins) {
        APIRL_HIP(c, hipStreamSynchronize(c->stream));
        (void)hipFree(p->d_divbins);
      }
      p->d_divbins = nullptr;
      p->divbins_cap = 0;
      if (hipMalloc(&p->d_divbins, nb * sizeof(float)) != hipSuccess)
        return fail(c, APIRL_ERR_NOMEM, "cannot allocate the divide-and-backproject bins");
      p->divbins_cap = nb;
    }
    APIRL_HIP(c, apirl::launch_divide_rule(sino, est, p->d_divbins, nb, p->geo.dims, c->stream));
    sino = p->d_divbins;
    kind = apirl::BP_PLAIN;
  }
  if (!ensure_tile_map(p))
    return fail(c, APIRL_ERR_NOMEM, "cannot allocate the gather's tile map");
  if (precision == APIRL_PREC_F64_SETUP && p->geo.dims == 3 && p->gather_zlanes <= 1)
    (void)ensure_f64s_tables(p);  // without them the one-voxel F64_SETUP kernel runs
  if (mode == APIRL_BP_TILE) {
    if (!ensure_tile_order(p))
      return fail(c, APIRL_ERR_NOMEM, "cannot allocate the tile backprojector's item order");
    if (tg && tg->ready && tg->stream != c->stream)
      APIRL_HIP(c, hipStreamWaitEvent(c->stream, tg->ev, 0));
    APIRL_HIP(c, apirl::launch_backproject_tile(p, sino, img, subset, nsubsets, kind,
                                                p->d_gscratch, p->gscratch_bytes, c->stream,
                                                fuse, tg ? tg->d : nullptr, tg && tg->ready,
                                                tg && tg->lists == 1 ? tg->list : nullptr));
    if (tg && !tg->ready) {
      tg->ready = true;
      tg->stream = c->stream;
      // the sinogram lists of its (item, ray) pairs: counted, then kept if they fit
      long long* dtot = nullptr;
      long long total = 0;
      APIRL_HIP(c, apirl::launch_tile_lists(p, subset, nsubsets, tg->d, nullptr, &dtot,
                                            c->stream));
      APIRL_HIP(c, hipMemcpyAsync(&total, dtot, sizeof(total), hipMemcpyDeviceToHost,
                                  c->stream));
      APIRL_HIP(c, hipStreamSynchronize(c->stream));
      const size_t lb = (size_t)(total > 0 ? total : 0) * sizeof(unsigned short);
      tg->lists = -1;
      if (lb > 0 && p->tile_geo_used + lb <= p->tile_geo_budget) {
        if (hipMalloc(&tg->list, lb) == hipSuccess) {
          p->tile_geo_used += lb;
          tg->list_bytes = lb;
          tg->lists = 1;
          APIRL_HIP(c, apirl::launch_tile_lists(p, subset, nsubsets, tg->d, tg->list, nullptr,
                                                c->stream));
        } else {
          (void)hipGetLastError();
          tg->list = nullptr;
        }
      }
      APIRL_HIP(c, hipEventRecord(tg->ev, c->stream));
    }
    return APIRL_OK;
  }
  APIRL_HIP(c, apirl::launch_backproject_gather(p, sino, img, subset, nsubsets, kind, precision,
                                                p->d_gscratch, p->gscratch_bytes, c->stream,
                                                fuse));
  return APIRL_OK;
}

}  // namespace

extern "C" {

// Per-device address of the allocation-free entry point's partials array, taken once
// per device with that device current.
double* noalloc_scratch_for(int device) {
  static std::mutex mu;
  static std::map<int, double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  double* p = nullptr;
  if (apirl::noalloc_scratch_address(&p) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;  // the entry point then reports the failure; not cached, so retried
    return p;
  }
  addr[device] = p;
  return p;
}

apirl_status apirl_ctx_create(int device, apirl_ctx** out) {
  if (!out) return APIRL_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return APIRL_ERR_HIP;
  if (device < 0 || device >= n) return APIRL_ERR_INVALID;
  apirl_ctx* c = new apirl_ctx();
  c->device = device;
  DeviceGuard g(device);
  if (hipMalloc(&c->d_fred, 64 * sizeof(float)) != hipSuccess ||
      hipMalloc(&c->d_ured, 8 * sizeof(unsigned long long)) != hipSuccess) {
    delete c;
    return APIRL_ERR_NOMEM;
  }
  c->d_noalloc = noalloc_scratch_for(device);
  *out = c;
  return APIRL_OK;
}

apirl_status apirl_ctx_destroy(apirl_ctx* ctx) {
  if (!ctx) return APIRL_OK;
  DeviceGuard g(ctx->device);
  if (ctx->d_red) (void)hipFree(ctx->d_red);
  if (ctx->d_fred) (void)hipFree(ctx->d_fred);
  if (ctx->d_ured) (void)hipFree(ctx->d_ured);
  delete ctx;
  return APIRL_OK;
}

const char* apirl_last_error(const apirl_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

apirl_status apirl_ctx_set_stream(apirl_ctx* ctx, void* stream) {
  if (!ctx) return APIRL_ERR_INVALID;
  ctx->stream = (hipStream_t)stream;
  return APIRL_OK;
}

apirl_status apirl_ctx_synchronize(apirl_ctx* ctx) {
  if (!ctx) return APIRL_ERR_INVALID;
  DeviceGuard g(ctx->device);
  APIRL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return APIRL_OK;
}

apirl_status apirl_malloc(apirl_ctx* ctx, void** dptr, size_t bytes) {
  if (!ctx || !dptr) return APIRL_ERR_INVALID;
  DeviceGuard g(ctx->device);
  if (hipMalloc(dptr, bytes ? bytes : 4) != hipSuccess)
    return fail(ctx, APIRL_ERR_NOMEM, "hipMalloc failed");
  return APIRL_OK;
}

apirl_status apirl_free(apirl_ctx* ctx, void* dptr) {
  if (!ctx) return APIRL_ERR_INVALID;
  DeviceGuard g(ctx->device);
  APIRL_HIP(ctx, hipFree(dptr));
  return APIRL_OK;
}

apirl_status apirl_memcpy_h2d(apirl_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return APIRL_ERR_INVALID;
  DeviceGuard g(ctx->device);
  APIRL_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  APIRL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return APIRL_OK;
}

apirl_status apirl_memcpy_d2h(apirl_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return APIRL_ERR_INVALID;
  DeviceGuard g(ctx->device);
  APIRL_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  APIRL_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return APIRL_OK;
}

apirl_status apirl_memset(apirl_ctx* ctx, void* dptr, int value, size_t bytes) {
  if (!ctx) return APIRL_ERR_INVALID;
  DeviceGuard g(ctx->device);
  APIRL_HIP(ctx, hipMemsetAsync(dptr, value, bytes, ctx->stream));
  return APIRL_OK;
}

namespace {
// apirl_plan_create; zab: the axial table to use instead of the one built from the
// mean ring coordinates (the parts of an oversampled plan)
apirl_status plan_create_impl(apirl_ctx* ctx, const apirl_image_desc* image,
                              const apirl_sino_desc* sino, const std::vector<float2>* zab,
                              apirl_plan** out) {
  if (!ctx || !image || !sino || !out) return APIRL_ERR_INVALID;
  *out = nullptr;
  const apirl_image_desc& im = *image;
  const apirl_sino_desc& sd = *sino;
  if (im.nx <= 0 || im.ny <= 0 || im.nz <= 0 || !(im.dx > 0) || !(im.dy > 0) || !(im.dz > 0))
    return fail(ctx, APIRL_ERR_INVALID, "invalid image size");
  if (sd.dims != 2 && sd.dims != 3) return fail(ctx, APIRL_ERR_INVALID, "dims must be 2 or 3");
  if (sd.nang <= 0 || sd.nr <= 0 || sd.nsino <= 0)
    return fail(ctx, APIRL_ERR_INVALID, "invalid sinogram size");
  if (sd.dims == 2 && (sd.nsino != 1 || im.nz != 1))
    return fail(ctx, APIRL_ERR_INVALID, "2D plans take one sinogram and a one-slice image");
  if (sd.dims == 3 && (!sd.z1_mean_mm || !sd.z2_mean_mm || !(sd.ring_width_mm > 0)))
    return fail(ctx, APIRL_ERR_INVALID, "3D plans need ring coordinates and ring width");
  if (!(sd.rscanner_mm > 0)) return fail(ctx, APIRL_ERR_INVALID, "invalid scanner radius");
  if ((double)im.nx * im.ny * im.nz > 2147483647.0 ||
      (double)sd.nsino * sd.nang * sd.nr > 9.0e15)
    return fail(ctx, APIRL_ERR_INVALID, "problem too large");
  DeviceGuard g(ctx->device);
  apirl_plan* p = new apirl_plan();
  p->ctx = ctx;
  p->img = im;
  p->sino = sd;
  p->sino.ang_deg = p->sino.r_mm = p->sino.z1_mean_mm = p->sino.z2_mean_mm = nullptr;
  p->ang.resize(sd.nang);
  p->r.resize(sd.nr);
  {
    std::vector<float> ga(sd.nang), gr(sd.nr);
    apirl_sino_grid(sd.nang, sd.nr, sd.rfov_mm, ga.data(), gr.data());
    for (int i = 0; i < sd.nang; i++) p->ang[i] = sd.ang_deg ? sd.ang_deg[i] : ga[i];
    for (int j = 0; j < sd.nr; j++) p->r[j] = sd.r_mm ? sd.r_mm[j] : gr[j];
  }
  std::vector<apirl::RayXY> rays;
  apirl::build_ray_table(im, sd, p->ang, p->r, rays);
  Geo& geo = p->geo;
  geo.nx = im.nx; geo.ny = im.ny; geo.nz = im.nz;
  geo.dx = im.dx; geo.dy = im.dy; geo.dz = im.dz;
  geo.R = (im.dx * im.nx) / 2;
  geo.offz = im.dz / 2;
  geo.nr = sd.nr; geo.nang = sd.nang; geo.nsino = sd.nsino;
  geo.dims = sd.dims;
  geo.ns = geo.nas = 1;
  {
    const long nxy = (long)im.nx * im.ny, lim = 1L << 23;
    geo.walk24 = (nxy < lim && im.nx < lim && im.ny < lim && im.nz < lim) ? 1 : 0;
  }
  if (hipMalloc(&p->d_rays, rays.size() * sizeof(apirl::RayXY)) != hipSuccess) {
    delete p;
    return fail(ctx, APIRL_ERR_NOMEM, "cannot allocate the LOR table");
  }
  if (hipMemcpy(p->d_rays, rays.data(), rays.size() * sizeof(apirl::RayXY),
                hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(p->d_rays);
    delete p;
    return fail(ctx, APIRL_ERR_HIP, "cannot upload the LOR table");
  }
  if (sd.dims == 3) {
    p->z1m.assign(sd.z1_mean_mm, sd.z1_mean_mm + sd.nsino);
    p->z2m.assign(sd.z2_mean_mm, sd.z2_mean_mm + sd.nsino);
    if (zab) p->zab = *zab;
    else apirl::build_z_table(sd, p->zab);
    if (hipMalloc(&p->d_zab, p->zab.size() * sizeof(float2)) != hipSuccess ||
        hipMemcpy(p->d_zab, p->zab.data(), p->zab.size() * sizeof(float2),
                  hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(p->d_rays);
      if (p->d_zab) (void)hipFree(p->d_zab);
      delete p;
      return fail(ctx, APIRL_ERR_NOMEM, "cannot upload the ring table");
    }
  }
  // Forward-projection order: sinograms by mean axial coordinate (stable).
  if (sd.dims == 3) {
    std::vector<int> order(sd.nsino);
    for (int q = 0; q < sd.nsino; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      return (p->zab[x].x + p->zab[x].y) < (p->zab[y].x + p->zab[y].y);
    });
    if (hipMalloc(&p->d_zorder, order.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(p->d_zorder, order.data(), order.size() * sizeof(int),
                  hipMemcpyHostToDevice) != hipSuccess) {
      if (p->d_zorder) (void)hipFree(p->d_zorder);
  if (p->d_imgT) (void)hipFree(p->d_imgT);
      p->d_zorder = nullptr;
      (void)hipGetLastError();
    }
  }
  // Tables of the gather backprojector's candidate search.
  {
    std::vector<float2> angcs;
    std::vector<apirl::GatherSeg> segs;
    float r0 = 0, dr = 1, rm = 0;
    bool ok = false;
    apirl::build_gather_tables(*p, angcs, segs, r0, dr, rm, ok);
    geo.r0 = r0;
    geo.dr = dr;
    geo.rmargin = rm;
    p->nseg = (int)segs.size();
    // 3D radial grids that are increasing but not uniform (explicit r_mm, e.g. the
    // mMR's arc-corrected positions) take the radial lookup search instead of the
    // closed form
    bool increasing = sd.nr >= 1;
    for (int j = 1; j < sd.nr; j++) increasing = increasing && p->r[j] > p->r[j - 1];
    p->uniform_r = ok;
    if (!ok && sd.dims == 3 && increasing && !segs.empty()) ok = true;
    p->gather_ok = ok;
    if (hipMalloc(&p->d_angcs, angcs.size() * sizeof(float2)) != hipSuccess ||
        hipMemcpy(p->d_angcs, angcs.data(), angcs.size() * sizeof(float2),
                  hipMemcpyHostToDevice) != hipSuccess)
      p->gather_ok = false;
    if (!segs.empty() &&
        (hipMalloc(&p->d_segs, segs.size() * sizeof(apirl::GatherSeg)) != hipSuccess ||
         hipMemcpy(p->d_segs, segs.data(), segs.size() * sizeof(apirl::GatherSeg),
                   hipMemcpyHostToDevice) != hipSuccess))
      p->gather_ok = false;
    (void)hipGetLastError();
  }
  // Slots of the gather's z-crossing table, indexed by slice: sinogram s of a
  // (angle, radial bin) block gets one entry per slice its LORs can enter (the slices
  // between its two ring planes, one more on each side for rounding), holding the
  // accumulated z crossing at which the ray enters that slice.
  if (sd.dims == 3 && p->gather_ok) {
    const int ns = (int)p->zab.size();
    std::vector<int> tab(2 * ns + 1, 0);  // [soff (ns + 1) | kmin (ns)]
    for (int s = 0; s < ns; s++) {
      const double z0 = std::min(p->zab[s].x, p->zab[s].y);
      const double z1 = std::max(p->zab[s].x, p->zab[s].y);
      const double offz = im.dz / 2.0;
      const int kmin = (int)std::floor((z0 - offz) / im.dz) - 1;
      const int kmax = (int)std::floor((z1 - offz) / im.dz) + 1;
      tab[s + 1] = tab[s] + (kmax - kmin + 1);
      tab[ns + 1 + s] = kmin;
    }
    p->zblock = tab[ns];
    if (hipMalloc(&p->d_zsoff, tab.size() * sizeof(int)) != hipSuccess ||
        hipMemcpy(p->d_zsoff, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess) {
      p->gather_ok = false;
      (void)hipGetLastError();
    }
  }
  // The plan's own LORs are form APIRL_SCANNER_CYLINDRICAL: the ray table uploaded
  // above, and for the gather its entry-crossing table (independent of subset and
  // sinogram) and, on a radial grid that is not uniform, the radial lookup table.
  {
    apirl::LorForm& f = p->forms[APIRL_SCANNER_CYLINDRICAL];
    f.rays = p->d_rays;
    f.built = true;
    p->gather_tables = p->gather_ok;  // candidate-search tables common to every form
    if (p->gather_tables) f.gather = upload_gather_form(p, rays, p->r, !p->uniform_r, false, f);
    apirl::use_lor_form(p, APIRL_SCANNER_CYLINDRICAL);
  }
  // The gather's tile order for the default depth (without it the column kernel
  // keeps the grid's order).
  (void)ensure_tile_map(p);
  // Transposed image for the forward projector's x-major rays.
  if (hipMalloc(&p->d_imgT, (size_t)im.nx * im.ny * im.nz * sizeof(float)) != hipSuccess) {
    p->d_imgT = nullptr;
    (void)hipGetLastError();
  }
  *out = p;
  return APIRL_OK;
}
}  // namespace

apirl_status apirl_plan_create(apirl_ctx* ctx, const apirl_image_desc* image,
                               const apirl_sino_desc* sino, apirl_plan** out) {
  return plan_create_impl(ctx, image, sino, nullptr, out);
}

extern "C++" {
namespace apirl {

void free_os_parts(apirl_plan* p) {
  for (apirl_plan* h : p->os_parts) apirl_plan_destroy(h);
  p->os_parts.clear();
}

bool ensure_os_parts(apirl_plan* p) {
  const int ns = p->geo.ns, nas = p->geo.nas;
  if (ns * nas == 1 || !p->os_parts.empty()) return true;
  apirl_ctx* c = p->ctx;
  // the LORs of the oversampled tables (apirl_plan_set_oversampling): radial sub-LOR o
  // of bin j at r_sub[j ns + o], axial sample q of sinogram s at zab[s nas + q]
  std::vector<float> r_sub;
  apirl::build_subsample_r(p->r, ns, r_sub);
  std::vector<float2> zab_os;
  apirl::build_z_table_os(p->z1m, p->z2m, p->sino.ring_width_mm, nas, zab_os);
  const int nr = p->sino.nr, nsino = p->sino.nsino;
  std::vector<float> r_o(nr), z1(nsino), z2(nsino);
  std::vector<float2> zab_q(nsino);
  for (int o = 0; o < ns; o++)
    for (int q = 0; q < nas; q++) {
      for (int j = 0; j < nr; j++) r_o[j] = r_sub[(size_t)j * ns + o];
      for (int s = 0; s < nsino; s++) {
        zab_q[s] = zab_os[(size_t)s * nas + q];
        z1[s] = zab_q[s].x;
        z2[s] = zab_q[s].y;
      }
      apirl_sino_desc sd = p->sino;
      sd.ang_deg = p->ang.data();
      sd.r_mm = r_o.data();
      sd.z1_mean_mm = z1.data();
      sd.z2_mean_mm = z2.data();
      apirl_plan* h = nullptr;
      if (plan_create_impl(c, &p->img, &sd, &zab_q, &h) != APIRL_OK || !h->gather_ok) {
        if (h) {
          apirl_plan_destroy(h);
          fail(c, APIRL_ERR_UNSUPPORTED, "the gather cannot search the oversampled LORs' grid");
        }
        free_os_parts(p);
        return false;
      }
      // a part only gathers: the forward projector's transposed image is not needed
      if (h->d_imgT) {
        (void)hipFree(h->d_imgT);
        h->d_imgT = nullptr;
      }
      h->gather_zvox = p->gather_zvox;
      h->gather_zlanes = p->gather_zlanes;
      h->gather_xcd = p->gather_xcd;
      h->tile_zs = p->tile_zs;
      h->tile_geo_budget = p->tile_geo_budget / ((size_t)ns * nas);  // the plan's budget
      (void)ensure_tile_map(h);
      p->os_parts.push_back(h);
    }
  return true;
}

}  // namespace apirl
}  // extern "C++"

apirl_status apirl_plan_destroy(apirl_plan* p) {
  if (!p) return APIRL_OK;
  DeviceGuard g(p->ctx->device);
  for (apirl::LorForm& f : p->forms) free_form(f);  // the active tables alias one
  if (p->d_zab) (void)hipFree(p->d_zab);
  if (p->d_replicas) (void)hipFree(p->d_replicas);
  apirl::free_os_parts(p);
  if (p->d_angcs) (void)hipFree(p->d_angcs);
  if (p->d_zorder) (void)hipFree(p->d_zorder);
  if (p->d_imgT) (void)hipFree(p->d_imgT);
  if (p->d_segs) (void)hipFree(p->d_segs);
  if (p->d_gscratch) (void)hipFree(p->d_gscratch);
  if (p->d_divbins) (void)hipFree(p->d_divbins);
  for (int* m : p->d_tilemap)
    if (m) (void)hipFree(m);
  if (p->d_tile_order) (void)hipFree(p->d_tile_order);
  free_tile_geo(p);
  if (p->d_zsoff) (void)hipFree(p->d_zsoff);
  delete p;
  return APIRL_OK;
}

size_t apirl_plan_subset_bins(const apirl_plan* p, int subset, int nsubsets) {
  if (!p || !check_subset(p, subset, nsubsets)) return 0;
  return (size_t)p->geo.nsino * apirl::subset_num_angles(p->geo.nang, nsubsets, subset) *
         p->geo.nr;
}

size_t apirl_plan_image_voxels(const apirl_plan* p) {
  return p ? (size_t)p->geo.nx * p->geo.ny * p->geo.nz : 0;
}

apirl_status apirl_plan_intersections(apirl_plan* p, int subset, int nsubsets,
                                      unsigned long long* out) {
  if (!p || !out || !check_subset(p, subset, nsubsets)) return APIRL_ERR_INVALID;
  apirl_ctx* c = p->ctx;
  DeviceGuard g(c->device);
  APIRL_HIP(c, hipMemsetAsync(c->d_ured, 0, sizeof(unsigned long long), c->stream));
  APIRL_HIP(c, apirl::launch_count_intersections(p, subset, nsubsets, c->d_ured, c->stream));
  APIRL_HIP(c, hipMemcpyAsync(out, c->d_ured, sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
  APIRL_HIP(c, hipStreamSynchronize(c->stream));
  return APIRL_OK;
}

apirl_status apirl_plan_set_oversampling(apirl_plan* p, int ns, int nas) {
  if (!p) return APIRL_ERR_INVALID;
  apirl_ctx* c = p->ctx;
  if (p->geo.dims != 3)
    return fail(c, APIRL_ERR_UNSUPPORTED, "oversampled LORs exist for 3D sinograms only");
  if (ns < 1 || nas < 1 || ns > 16 || nas > 16 || ns * nas > 64)
    return fail(c, APIRL_ERR_INVALID, "samples per bin must be 1..16 each, 64 in all");
  if (ns == p->geo.ns && nas == p->geo.nas) return APIRL_OK;
  if (p->scanner != APIRL_SCANNER_CYLINDRICAL)
    return fail(c, APIRL_ERR_UNSUPPORTED, "oversampled LORs are defined for the cylindrical scanner");
  DeviceGuard g(c->device);
  // radial sub-LORs: the bin's own LOR when ns == 1 (getPointsFromLor), the
  // subsampled r otherwise (getPointsFromOverSampledLor)
  std::vector<float> r_sub;
  if (ns == 1) r_sub = p->r;
  else apirl::build_subsample_r(p->r, ns, r_sub);
  apirl_sino_desc sd = p->sino;
  sd.nr = p->sino.nr * ns;
  std::vector<apirl::RayXY> rays;
  apirl::build_ray_table(p->img, sd, p->ang, r_sub, rays);
  std::vector<float2> zab;
  apirl::build_z_table_os(p->z1m, p->z2m, p->sino.ring_width_mm, nas, zab);
  apirl::RayXY* d_rays = nullptr;
  float2* d_zab = nullptr;
  if (hipMalloc(&d_rays, rays.size() * sizeof(apirl::RayXY)) != hipSuccess ||
      hipMalloc(&d_zab, zab.size() * sizeof(float2)) != hipSuccess) {
    if (d_rays) (void)hipFree(d_rays);
    (void)hipGetLastError();
    return fail(c, APIRL_ERR_NOMEM, "cannot allocate the oversampled LOR tables");
  }
  if (hipMemcpy(d_rays, rays.data(), rays.size() * sizeof(apirl::RayXY),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_zab, zab.data(), zab.size() * sizeof(float2), hipMemcpyHostToDevice) !=
          hipSuccess) {
    (void)hipFree(d_rays);
    (void)hipFree(d_zab);
    return fail(c, APIRL_ERR_HIP, "cannot upload the oversampled LOR tables");
  }
  APIRL_HIP(c, hipStreamSynchronize(c->stream));
  apirl::free_os_parts(p);
  apirl::LorForm& f = p->forms[APIRL_SCANNER_CYLINDRICAL];
  free_f64s(f);  // built from the rays replaced here
  (void)hipFree(f.rays);
  (void)hipFree(p->d_zab);
  f.rays = d_rays;
  p->d_zab = d_zab;
  p->geo.ns = ns;
  p->geo.nas = nas;
  apirl::use_lor_form(p, APIRL_SCANNER_CYLINDRICAL);  // the gather needs one LOR per bin
  return APIRL_OK;
}

apirl_status apirl_plan_set_scanner(apirl_plan* p, int scanner) {
  if (!p) return APIRL_ERR_INVALID;
  apirl_ctx* c = p->ctx;
  if (scanner != APIRL_SCANNER_CYLINDRICAL && scanner != APIRL_SCANNER_MMR &&
      scanner != APIRL_SCANNER_MMR_SUBSET)
    return fail(c, APIRL_ERR_INVALID, "unknown scanner");
  if (scanner == p->scanner) return APIRL_OK;
  if (p->geo.dims != 3)
    return fail(c, APIRL_ERR_UNSUPPORTED, "scanner geometries are set on 3D plans");
  if (p->geo.ns * p->geo.nas != 1)
    return fail(c, APIRL_ERR_UNSUPPORTED, "set the scanner before oversampling");
  DeviceGuard g(c->device);
  if (!apirl::build_lor_form(p, scanner)) return APIRL_ERR_NOMEM;
  apirl::use_lor_form(p, scanner);
  return APIRL_OK;
}

apirl_status apirl_plan_get_tuning(const apirl_plan* p, apirl_tuning* out) {
  if (!p || !out) return APIRL_ERR_INVALID;
  out->gather_zlanes = p->gather_zlanes;
  out->gather_zvox = p->gather_zvox;
  out->gather_xcd = p->gather_xcd;
  out->tile_zs = p->tile_zs;
  out->tile_cache_mb = (int)(p->tile_geo_budget >> 20);
  return APIRL_OK;
}

apirl_status apirl_plan_set_tuning(apirl_plan* p, const apirl_tuning* t) {
  if (!p || !t) return APIRL_ERR_INVALID;
  if (t->gather_zlanes) {
    if (t->gather_zlanes < 1 || t->gather_zlanes > 64)
      return fail(p->ctx, APIRL_ERR_INVALID, "gather_zlanes must be in 1..64");
    p->gather_zlanes = t->gather_zlanes;
  }
  if (t->gather_zvox) {
    if (t->gather_zvox < 1 || t->gather_zvox > 4)
      return fail(p->ctx, APIRL_ERR_INVALID, "gather_zvox must be in 1..4");
    p->gather_zvox = t->gather_zvox;
  }
  if (t->gather_xcd) {
    if (t->gather_xcd != APIRL_GATHER_XCD_ROUND_ROBIN && t->gather_xcd != APIRL_GATHER_XCD_WEDGES)
      return fail(p->ctx, APIRL_ERR_INVALID, "unknown gather_xcd");
    p->gather_xcd = t->gather_xcd;
  }
  bool zs_changed = false;
  if (t->tile_zs) {
    if (t->tile_zs < 1 || t->tile_zs > 256)
      return fail(p->ctx, APIRL_ERR_INVALID, "tile_zs must be in 1..256");
    zs_changed = t->tile_zs != p->tile_zs;
    p->tile_zs = t->tile_zs;
  }
  DeviceGuard g(p->ctx->device);
  if (t->tile_cache_mb) {
    if (t->tile_cache_mb < -1)
      return fail(p->ctx, APIRL_ERR_INVALID, "tile_cache_mb must be positive, or -1 for none");
    p->tile_geo_budget = t->tile_cache_mb < 0 ? 0 : (size_t)t->tile_cache_mb << 20;
    if (p->tile_geo_used > p->tile_geo_budget) free_tile_geo(p);
  }
  if (zs_changed) free_tile_geo(p);  // the kept lists are per slab
  if (!ensure_tile_map(p))
    return fail(p->ctx, APIRL_ERR_NOMEM, "cannot allocate the gather's tile map");
  for (apirl_plan* h : p->os_parts) {  // the oversampled plan's gathers follow it
    h->gather_zlanes = p->gather_zlanes;
    h->gather_zvox = p->gather_zvox;
    h->gather_xcd = p->gather_xcd;
    h->tile_zs = p->tile_zs;
    if (zs_changed) free_tile_geo(h);
    const size_t share = p->tile_geo_budget / p->os_parts.size();
    if (h->tile_geo_budget != share) {
      h->tile_geo_budget = share;
      if (h->tile_geo_used > h->tile_geo_budget) free_tile_geo(h);
    }
    if (!ensure_tile_map(h))
      return fail(p->ctx, APIRL_ERR_NOMEM, "cannot allocate the gather's tile map");
  }
  return APIRL_OK;
}

apirl_status apirl_project(apirl_plan* p, const float* d_img, float* d_sino, int subset,
                           int nsubsets, int precision) {
  if (!p || !d_img || !d_sino) return APIRL_ERR_INVALID;
  if (!check_subset(p, subset, nsubsets))
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid subset");
  if (precision < APIRL_PREC_REFERENCE || precision > APIRL_PREC_F64_SETUP)
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid precision");
  if (!os_precision_ok(p, precision))
    return fail(p->ctx, APIRL_ERR_UNSUPPORTED, "oversampled LORs use the reference precision");
  DeviceGuard g(p->ctx->device);
  APIRL_HIP(p->ctx, apirl::launch_project(p, d_img, d_sino, subset, nsubsets, precision,
                                          nullptr, nullptr, p->ctx->stream));
  return APIRL_OK;
}

apirl_status apirl_backproject(apirl_plan* p, const float* d_sino, float* d_img, int subset,
                               int nsubsets, int mode, int precision) {
  if (!p || !d_img || !d_sino) return APIRL_ERR_INVALID;
  if (!check_subset(p, subset, nsubsets))
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid subset");
  if (mode != APIRL_BP_SCATTER && mode != APIRL_BP_GATHER && mode != APIRL_BP_TILE)
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid backprojection mode");
  DeviceGuard g(p->ctx->device);
  return backproject_any(p, d_sino, nullptr, d_img, subset, nsubsets, apirl::BP_PLAIN, mode,
                         precision);
}

apirl_status apirl_backproject_update(apirl_plan* p, const float* d_sino, float* d_img,
                                      const float* d_sens, float thr, int strict, int subset,
                                      int nsubsets, int mode, int precision) {
  if (!p || !d_img || !d_sino || !d_sens) return APIRL_ERR_INVALID;
  if (!check_subset(p, subset, nsubsets))
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid subset");
  DeviceGuard g(p->ctx->device);
  if (mode != APIRL_BP_GATHER && mode != APIRL_BP_TILE)
    return fail(p->ctx, APIRL_ERR_INVALID, "the fused update takes the gather or tile form");
  const apirl::EmFuse fz{d_sens, thr, strict ? 1 : 0};
  return backproject_any(p, d_sino, nullptr, d_img, subset, nsubsets, apirl::BP_PLAIN, mode,
                         precision, &fz);
}

apirl_status apirl_divide_backproject(apirl_plan* p, const float* d_meas, const float* d_est,
                                      float* d_img, int subset, int nsubsets, int mode) {
  if (!p || !d_meas || !d_est || !d_img) return APIRL_ERR_INVALID;
  if (!check_subset(p, subset, nsubsets))
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid subset");
  if (mode != APIRL_BP_SCATTER && mode != APIRL_BP_GATHER && mode != APIRL_BP_TILE)
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid backprojection mode");
  DeviceGuard g(p->ctx->device);
  return backproject_any(p, d_meas, d_est, d_img, subset, nsubsets, apirl::BP_DIVIDE, mode);
}

namespace {
// The fused forward step.  noalloc: the likelihood partials go to the file-scope
// array of this GPU (no allocation; up to kNoallocPartials blocks), otherwise to the
// context's buffer, grown on demand.
apirl_status project_ratio_impl(apirl_plan* p, const float* d_img, const float* d_meas,
                                const float* d_additive, float* d_ratio, double* d_lik,
                                int subset, int nsubsets, int flags, int precision,
                                bool noalloc) {
  if (!p || !d_img || !d_meas || !d_ratio) return APIRL_ERR_INVALID;
  if (!check_subset(p, subset, nsubsets))
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid subset");
  apirl_ctx* c = p->ctx;
  if (precision < APIRL_PREC_REFERENCE || precision > APIRL_PREC_F64_SETUP)
    return fail(c, APIRL_ERR_INVALID, "invalid precision");
  if (!os_precision_ok(p, precision))
    return fail(c, APIRL_ERR_UNSUPPORTED, "oversampled LORs use the reference precision");
  DeviceGuard g(c->device);
  const int nblocks =
      apirl::project_partials(p, subset, nsubsets, apirl::path_forward_ok(p, precision));
  apirl::FusedRatio fr;
  fr.meas = d_meas;
  fr.additive = d_additive;
  fr.ratio = d_ratio;
  if (flags & ~(APIRL_RATIO_LIK_POSITIVE | APIRL_RATIO_ADDITIVE_LOCAL))
    return fail(c, APIRL_ERR_INVALID, "unknown project_ratio flags");
  fr.lik_positive_only = (flags & APIRL_RATIO_LIK_POSITIVE) ? 1 : 0;
  fr.additive_local = (flags & APIRL_RATIO_ADDITIVE_LOCAL) ? 1 : 0;
  if (d_lik) {
    if (noalloc) {
      // the documented limit in bins (one partial per 256 bins at most, whichever
      // projector kernel runs)
      if (apirl_plan_subset_bins(p, subset, nsubsets) > (size_t)APIRL_NOALLOC_MAX_BINS ||
          nblocks > apirl::kNoallocPartials)
        return fail(c, APIRL_ERR_INVALID,
                    "apirl_project_ratio_noalloc: the likelihood of more than "
                    "APIRL_NOALLOC_MAX_BINS bins per call needs apirl_project_ratio");
      if (!c->d_noalloc)
        return fail(c, APIRL_ERR_HIP, "the partials array's device address is unavailable");
      fr.lik_partials = c->d_noalloc;
    } else {
      apirl_status st = ensure_red(c, (size_t)nblocks);
      if (st != APIRL_OK) return st;
      fr.lik_partials = c->d_red;
    }
  }
  APIRL_HIP(c, apirl::launch_project(p, d_img, nullptr, subset, nsubsets, precision, &fr,
                                     nullptr, c->stream));
  if (d_lik)
    APIRL_HIP(c, apirl::launch_sum_partials(fr.lik_partials, nblocks, d_lik, c->stream));
  return APIRL_OK;
}
}  // namespace

apirl_status apirl_project_ratio(apirl_plan* p, const float* d_img, const float* d_meas,
                                 const float* d_additive, float* d_ratio, double* d_lik,
                                 int subset, int nsubsets, int flags, int precision) {
  return project_ratio_impl(p, d_img, d_meas, d_additive, d_ratio, d_lik, subset, nsubsets,
                            flags, precision, false);
}

apirl_status apirl_project_ratio_noalloc(apirl_plan* p, const float* d_img,
                                         const float* d_meas, const float* d_additive,
                                         float* d_ratio, double* d_lik, int subset,
                                         int nsubsets, int flags, int precision) {
  return project_ratio_impl(p, d_img, d_meas, d_additive, d_ratio, d_lik, subset, nsubsets,
                            flags, precision, true);
}

apirl_status apirl_em_update(apirl_ctx* c, float* d_x, const float* d_b, const float* d_s,
                             size_t n, float thr, int strict) {
  if (!c || !d_x || !d_b || !d_s) return APIRL_ERR_INVALID;
  DeviceGuard g(c->device);
  APIRL_HIP(c, apirl::launch_em_update(d_x, d_b, d_s, n, thr, strict, c->stream));
  return APIRL_OK;
}

namespace {
apirl_status sensitivity_impl(apirl_plan* p, const float* d_mult, float* d_sens, int subset,
                              int nsubsets, int mode, int force_positive, int precision) {
  if (!p || !d_sens) return APIRL_ERR_INVALID;
  if (!check_subset(p, subset, nsubsets))
    return fail(p->ctx, APIRL_ERR_INVALID, "invalid subset");
  apirl_ctx* c = p->ctx;
  DeviceGuard g(c->device);
  const size_t nvox = apirl_plan_image_voxels(p);
  APIRL_HIP(c, hipMemsetAsync(d_sens, 0, nvox * sizeof(float), c->stream));
  apirl_status st = backproject_any(p, d_mult, nullptr, d_sens, subset, nsubsets,
                                    d_mult ? apirl::BP_MULT : apirl::BP_ONES, mode, precision);
  if (st != APIRL_OK) return st;
  if (force_positive) APIRL_HIP(c, apirl::launch_force_positive(d_sens, nvox, c->stream));
  return APIRL_OK;
}
}  // namespace

apirl_status apirl_sensitivity(apirl_plan* p, const float* d_mult, float* d_sens, int subset,
                               int nsubsets, int mode, int force_positive) {
  return sensitivity_impl(p, d_mult, d_sens, subset, nsubsets, mode, force_positive,
                          APIRL_PREC_REFERENCE);
}

apirl_status apirl_mlem_threshold(apirl_ctx* c, const float* d_sens, size_t n, float* h_thr) {
  if (!c || !d_sens || !h_thr || n == 0) return APIRL_ERR_INVALID;
  DeviceGuard g(c->device);
  float mm[2];
  APIRL_HIP(c, apirl::launch_min_nz_max(d_sens, n, c->d_fred, c->stream));
  APIRL_HIP(c, hipMemcpyAsync(mm, c->d_fred, 2 * sizeof(float), hipMemcpyDeviceToHost,
                              c->stream));
  APIRL_HIP(c, hipStreamSynchronize(c->stream));
  // Mlem.cpp:41
  *h_thr = mm[0] + (mm[1] - mm[0]) * 0.0001f;
  return APIRL_OK;
}

apirl_status apirl_em_reconstruct(apirl_plan* p, const apirl_em_options* opt,
                                  const float* d_meas, const float* d_init,
                                  const float* d_additive, const float* d_mult, float* d_img,
                                  double* h_lik, float* d_sens_out) {
  return apirl_em_reconstruct_timed(p, opt, d_meas, d_init, d_additive, d_mult, d_img, h_lik,
                                    d_sens_out, nullptr);
}

apirl_status apirl_em_reconstruct_timed(apirl_plan* p, const apirl_em_options* opt,
                                        const float* d_meas, const float* d_init,
                                        const float* d_additive, const float* d_mult,
                                        float* d_img, double* h_lik, float* d_sens_out,
                                        float* h_times) {
  if (!p || !opt || !d_meas || !d_init || !d_img) return APIRL_ERR_INVALID;
  apirl_ctx* c = p->ctx;
  const int S = opt->subsets, T = opt->iterations;
  if (S < 1 || S > p->geo.nang || T < 0)
    return fail(c, APIRL_ERR_INVALID, "invalid iterations / subsets");
  DeviceGuard g(c->device);
  const bool is2d = p->geo.dims == 2;
  const size_t nvox = apirl_plan_image_voxels(p);
  const size_t nfull = apirl_plan_subset_bins(p, 0, 1);
  size_t nsub_max = 0;
  for (int s = 0; s < S; s++) {
    size_t n = apirl_plan_subset_bins(p, s, S);
    if (n > nsub_max) nsub_max = n;
  }
  // Device-resident state: one sensitivity image per subset (OsemSinogram3d.h:76),
  // the ratio sinogram of the largest subset and the backprojected image.
  float *d_sens = nullptr, *d_ratio = nullptr, *d_bp = nullptr;
  double* d_lik = nullptr;
  std::vector<hipEvent_t> ev;  // phase timing events (h_times)
  auto cleanup = [&]() {
    for (auto& x : ev)
      if (x) (void)hipEventDestroy(x);
    ev.clear();
    if (d_sens && d_sens != d_sens_out) (void)hipFree(d_sens);
    if (d_ratio) (void)hipFree(d_ratio);
    if (d_bp) (void)hipFree(d_bp);
    if (d_lik) (void)hipFree(d_lik);
  };
  if (d_sens_out) d_sens = d_sens_out;
  else if (hipMalloc(&d_sens, (size_t)S * nvox * sizeof(float)) != hipSuccess) d_sens = nullptr;
  if (!d_sens || hipMalloc(&d_ratio, std::max(nsub_max, nfull) * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_bp, nvox * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_lik, ((size_t)T * S + 1) * sizeof(double)) != hipSuccess) {
    cleanup();
    return fail(c, APIRL_ERR_NOMEM, "cannot allocate the reconstruction state");
  }
  apirl_status st = APIRL_OK;
#define APIRL_TRY(x)            \
  do {                          \
    st = (x);                   \
    if (st != APIRL_OK) {       \
      cleanup();                \
      return st;                \
    }                           \
  } while (0)
  // OSEM on the mMR: every subset sinogram the reference builds carries cylindrical
  // LORs (APIRL_SCANNER_MMR_SUBSET), sensitivities included; only the final full
  // projection keeps the mMR's (OsemSinogram3d.cpp:101-117,164-166,291-292).
  struct FormRestore {
    apirl_plan* p;
    int form;
    bool on;
    ~FormRestore() {
      if (on) apirl::use_lor_form(p, form);
    }
  } restore{p, p->scanner, false};
  if (p->scanner == APIRL_SCANNER_MMR && opt->osem_rules) {
    if (!apirl::build_lor_form(p, APIRL_SCANNER_MMR_SUBSET)) {
      cleanup();
      return APIRL_ERR_NOMEM;
    }
    apirl::use_lor_form(p, APIRL_SCANNER_MMR_SUBSET);
    restore.on = true;
  }
  // The precision mode applies to the forward projector and, for the gather form,
  // to the backprojector; the scatter form always uses the reference's arithmetic.
  const int bp_prec = opt->bp_mode == APIRL_BP_GATHER ? opt->precision : APIRL_PREC_REFERENCE;
  std::vector<float> thr(S, 0.f);
  if (opt->sensitivity_given && !d_sens_out) {
    cleanup();
    return fail(c, APIRL_ERR_INVALID, "sensitivity_given needs the images in d_sens_out");
  }
  for (int s = 0; s < S; s++) {
    if (!opt->sensitivity_given) {
      APIRL_TRY(sensitivity_impl(p, d_mult, d_sens + (size_t)s * nvox, s, S, opt->bp_mode,
                                 opt->osem_rules ? 1 : 0, bp_prec));
    } else if (opt->osem_rules) {
      // images read from files are clamped as computed ones (OsemSinogram3d.cpp:93)
      hipError_t e = apirl::launch_force_positive(d_sens + (size_t)s * nvox, nvox, c->stream);
      if (e != hipSuccess) {
        cleanup();
        return hip_fail(c, e, "forcePositive");
      }
    }
    if (!opt->osem_rules) APIRL_TRY(apirl_mlem_threshold(c, d_sens + (size_t)s * nvox, nvox, &thr[s]));
  }
  // MLEM: 2D updates with s > thr (Mlem2d.cpp:175), 3D with s >= thr
  // (MlemSinogram3d.cpp:231); OSEM: threshold 0 with >= (OsemSinogram3d.cpp:28,221).
  const int strict = (!opt->osem_rules && is2d) ? 1 : 0;
  // With the gather (one LOR per bin) the update is fused into the backprojection: the
  // same float backprojection per voxel, without storing it (bit-identical images).
  const bool fuse_update = (opt->bp_mode == APIRL_BP_GATHER || opt->bp_mode == APIRL_BP_TILE) &&
                           p->geo.ns * p->geo.nas == 1;
  // Likelihood rule (2D: est > 0) and the additive sinogram's angle per subset bin:
  // the reference adds angle k of the full additive sinogram to subset bin k
  // (OsemSinogram3d.cpp:172-173, Sinogram3D.cpp:857-873).
  const int ratio_flags = (is2d ? APIRL_RATIO_LIK_POSITIVE : 0) |
                          (opt->additive_own_angles ? 0 : APIRL_RATIO_ADDITIVE_LOCAL);
  if (hipMemcpyAsync(d_img, d_init, nvox * sizeof(float), hipMemcpyDeviceToDevice, c->stream) !=
      hipSuccess) {
    cleanup();
    return fail(c, APIRL_ERR_HIP, "image copy failed");
  }
  // Per sub-iteration phase timing (h_times): events before the projection, the
  // backprojection, the update and after it.
  if (h_times) {
    ev.assign((size_t)T * S * 3 + 1, nullptr);
    for (auto& x : ev)
      if (hipEventCreate(&x) != hipSuccess) {
        x = nullptr;
        cleanup();
        return fail(c, APIRL_ERR_HIP, "event creation failed");
      }
  }
  auto mark = [&](size_t i) {
    if (h_times) (void)hipEventRecord(ev[i], c->stream);
  };
  for (int t = 0; t < T; t++) {
    for (int s = 0; s < S; s++) {
      const size_t q = ((size_t)t * S + s) * 3;
      mark(q);
      APIRL_TRY(apirl_project_ratio(p, d_img, d_meas, d_additive, d_ratio,
                                    d_lik + (size_t)t * S + s, s, S, ratio_flags,
                                    opt->precision));
      mark(q + 1);
      if (fuse_update) {
        // the gather updates each voxel of the image from its backprojection
        const apirl::EmFuse fz{d_sens + (size_t)s * nvox, thr[s], strict};
        APIRL_TRY(backproject_any(p, d_ratio, nullptr, d_img, s, S, apirl::BP_PLAIN,
                                  opt->bp_mode, bp_prec, &fz));
        mark(q + 2);
        continue;
      }
      if (hipMemsetAsync(d_bp, 0, nvox * sizeof(float), c->stream) != hipSuccess) {
        cleanup();
        return fail(c, APIRL_ERR_HIP, "memset failed");
      }
      APIRL_TRY(apirl_backproject(p, d_ratio, d_bp, s, S, opt->bp_mode, bp_prec));
      mark(q + 2);
      APIRL_TRY(apirl_em_update(c, d_img, d_bp, d_sens + (size_t)s * nvox, nvox, thr[s], strict));
    }
  }
  mark((size_t)T * S * 3);
  // Final likelihood.  3D drivers project the final image once more over the full
  // sinogram, without the additive term (MlemSinogram3d.cpp:277-278,
  // OsemSinogram3d.cpp:291-293).  Mlem2d takes it on the last estimated sinogram,
  // which by then holds the ratio (Mlem2d.cpp:227-228).
  if (restore.on) {
    apirl::use_lor_form(p, restore.form);
    restore.on = false;
  }
  if (!is2d) {
    APIRL_TRY(apirl_project_ratio(p, d_img, d_meas, nullptr, d_ratio,
                                  d_lik + (size_t)T * S, 0, 1, 0, opt->precision));
  } else {
    const size_t n_last = apirl_plan_subset_bins(p, S - 1, S);
    APIRL_TRY(ensure_red(c, (n_last + 255) / 256));
    if (T == 0) {
      if (hipMemsetAsync(d_ratio, 0, n_last * sizeof(float), c->stream) != hipSuccess) {
        cleanup();
        return fail(c, APIRL_ERR_HIP, "memset failed");
      }
    }
    hipError_t e = apirl::launch_likelihood(d_ratio, d_meas, (long)n_last, 1, c->d_red,
                                            d_lik + (size_t)T * S, c->stream);
    if (e != hipSuccess) {
      cleanup();
      return hip_fail(c, e, "likelihood");
    }
  }
  if (h_lik) {
    if (hipMemcpyAsync(h_lik, d_lik, ((size_t)T * S + 1) * sizeof(double),
                       hipMemcpyDeviceToHost, c->stream) != hipSuccess) {
      cleanup();
      return fail(c, APIRL_ERR_HIP, "likelihood copy failed");
    }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    cleanup();
    return fail(c, APIRL_ERR_HIP, "stream synchronize failed");
  }
  if (h_times) {
    // [total | forward | backward | update] x (T S), in ms
    const size_t n = (size_t)T * S;
    for (size_t k = 0; k < n; k++) {
      float f = 0.f, b = 0.f, u = 0.f;
      (void)hipEventElapsedTime(&f, ev[3 * k], ev[3 * k + 1]);
      (void)hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]);
      (void)hipEventElapsedTime(&u, ev[3 * k + 2], ev[3 * k + 3]);
      h_times[k] = f + b + u;
      h_times[n + k] = f;
      h_times[2 * n + k] = b;
      h_times[3 * n + k] = u;
    }
  }
#undef APIRL_TRY
  cleanup();
  return APIRL_OK;
}

}  // extern "C"

namespace apirl {

int use_lor_form(apirl_plan* p, int f) {
  const int prev = p->scanner;
  const LorForm& L = p->forms[f];
  p->d_rays = L.rays;
  p->d_xytab = L.xytab;
  p->d_angcs_odd = L.angcs_odd;
  p->d_rlut = L.rlut;
  p->rl0 = L.rl0;
  p->rl_inv_h = L.rl_inv_h;
  p->nrl = L.nrl;
  p->scanner = f;
  // the gather's candidate search and tables assume one LOR per bin
  p->gather_ok = p->gather_tables && L.gather && p->geo.ns == 1 && p->geo.nas == 1;
  return prev;
}

bool build_lor_form(apirl_plan* p, int form) {
  LorForm& f = p->forms[form];
  if (f.built) return true;
  apirl_ctx* c = p->ctx;
  if (p->geo.dims != 3) {
    fail(c, APIRL_ERR_UNSUPPORTED, "scanner geometries are set on 3D plans");
    return false;
  }
  std::vector<RayXY> rays;
  std::vector<float> r;
  if (form == APIRL_SCANNER_CYLINDRICAL) {
    build_ray_table(p->img, p->sino, p->ang, p->r, rays);
    r = p->r;
  } else {
    std::vector<float4> pts;
    if (form == APIRL_SCANNER_MMR) mmr_lor_points(p->geo.nang, p->geo.nr, pts);
    else mmr_subset_lor_points(p->geo.nang, p->geo.nr, pts);
    std::vector<float2> odd;
    mmr_radial_grid(p->geo.nang, p->geo.nr, r, odd);
    std::vector<char> ok(pts.size(), 1);  // the 3D form rejects no LOR
    build_ray_table_pts(p->img, false, pts, ok, rays);
  }
  if (hipMalloc(&f.rays, rays.size() * sizeof(RayXY)) != hipSuccess ||
      hipMemcpy(f.rays, rays.data(), rays.size() * sizeof(RayXY), hipMemcpyHostToDevice) !=
          hipSuccess) {
    (void)hipGetLastError();
    free_form(f);
    fail(c, APIRL_ERR_NOMEM, "cannot upload the LOR table");
    return false;
  }
  if (p->gather_tables) {
    if (form == APIRL_SCANNER_CYLINDRICAL)
      f.gather = upload_gather_form(p, rays, p->r, !p->uniform_r, false, f);
    else f.gather = upload_gather_form(p, rays, r, true, form == APIRL_SCANNER_MMR, f);
  }
  f.built = true;
  return true;
}

}  // namespace apirl
