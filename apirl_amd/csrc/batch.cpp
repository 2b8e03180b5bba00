// Batched OSEM / MLEM jobs: many reconstructions of different geometries advance
// together, one sub-iteration of every active job per round.  Jobs of at least
// kLargeVoxels voxels run their sub-iteration with the single-job kernels after the
// grouped launches; the others share the round's seven grouped launches (clear, fused projection + ratio for 2D and 3D jobs, axial setup,
// gather backprojection for 2D and 3D jobs, update) that read per-job descriptors
// from a table in device memory; the launches are persistent (fixed grid), so the
// whole round is captured once into a hipGraph and replayed, and jobs join or leave
// between rounds by rewriting the table (continuous batching).
//
// Each job computes exactly what apirl_em_reconstruct computes for it with
// APIRL_BP_GATHER and APIRL_PREC_REFERENCE (same per-ray and per-voxel code, fixed
// summation order), so its image does not depend on which other jobs share the
// rounds.  The likelihood is not evaluated in batched rounds.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "apirl_internal.h"

namespace {

using namespace apirl;

struct BJob {
  int id = -1;
  apirl_plan* plan = nullptr;
  int T = 0, S = 1, osem = 1;
  bool is2d = false;
  const float* meas = nullptr;
  const float* add = nullptr;
  int add_local = 1;         // additive angle per subset bin (FusedRatio::additive_local)
  int form = 0;              // LOR form of the plan its sub-iterations use (APIRL_SCANNER_*)
  float* x = nullptr;
  size_t nvox = 0;
  int t = 0, s = 0;          // next sub-iteration to schedule
  bool active = false;       // scheduled in future rounds
  bool done_recorded = false;
  hipEvent_t done_ev = nullptr;
  std::vector<float> thr;
  int strict = 0;
  // device state owned by the job
  float* sens = nullptr;     // S images
  float* ratio = nullptr;    // largest subset sinogram
  float* bp = nullptr;       // backprojection
  void* zscr = nullptr;      // gather axial records + non-zero flags (3D)
  size_t zscr_bytes = 0;
  bool large = false;        // runs its sub-iterations with the single-job kernels
  bool tile = false;         //   and, large, the tile backprojector (else the gather)
  std::vector<std::pair<void*, size_t>> bufs;
};

// A 3D job of at least this many voxels fills the GPU with its own launches, so its
// sub-iterations are enqueued with the single-job kernels (which read the plan's
// transposed image copy and need no job table) after each grouped round.
constexpr size_t kLargeVoxels = size_t(1) << 21;

template <class T>
size_t align_up(size_t n) {
  return (n + 255) / 256 * 256;
}

}  // namespace

struct apirl_batch {
  apirl_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  int cap = 0, max_nsino = 0;
  int grid = 2048;
  bool use_graph = true;
  int grab = 2;                // virtual blocks per counter draw (apirl_batch_set_schedule)
  int order_mode = APIRL_BATCH_ORDER_SUBMISSION;
  std::vector<BJob*> jobs;         // by id
  std::vector<int> order;          // active job ids in admission order
  // job tables: one device copy, two pinned staging copies
  char* d_tab = nullptr;
  char* h_tab[2] = {nullptr, nullptr};
  hipEvent_t copy_ev[2] = {nullptr, nullptr};
  int cur = 0;
  size_t off_clear = 0, off_p2 = 0, off_p3 = 0, off_zs = 0, off_g2 = 0, off_g3 = 0,
         off_upd = 0, tab_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::multimap<size_t, void*> pool;  // released device buffers, reused by size
  std::string err;
};

namespace {

apirl_status bfail(apirl_batch* b, apirl_status st, const std::string& msg) {
  b->err = msg;
  if (b->ctx) b->ctx->err = msg;
  return st;
}

void* balloc(apirl_batch* b, size_t bytes) {
  auto it = b->pool.find(bytes);
  if (it != b->pool.end()) {
    void* p = it->second;
    b->pool.erase(it);
    return p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

void release_bufs(apirl_batch* b, BJob* j) {
  for (auto& pr : j->bufs) b->pool.emplace(pr.second, pr.first);
  j->bufs.clear();
  j->sens = j->ratio = j->bp = nullptr;
  j->zscr = nullptr;
}

size_t max_subset_bins(const apirl_plan* p, int S) {
  size_t m = 0;
  for (int s = 0; s < S; s++) m = std::max(m, apirl_plan_subset_bins(p, s, S));
  return m;
}

size_t max_gather_scratch(const apirl_plan* p, int S) {
  size_t m = 0;
  for (int s = 0; s < S; s++) m = std::max(m, gather_scratch_bytes(p, s, S));
  return m;
}

// The direct path's gather backprojection of a job's ratio into its bp: the plan's
// own, or for oversampled bins those of its plans of one LOR per bin, in turn.
hipError_t backproject_job(const BJob* j, int s, hipStream_t st) {
  const apirl_plan* p = j->plan;
  if (p->geo.ns * p->geo.nas == 1)
    return launch_backproject_gather(p, j->ratio, j->bp, s, j->S, BP_PLAIN, APIRL_PREC_REFERENCE,
                                     j->zscr, j->zscr_bytes, st);
  for (const apirl_plan* h : p->os_parts) {
    const hipError_t e = launch_backproject_gather(h, j->ratio, j->bp, s, j->S, BP_PLAIN,
                                                   APIRL_PREC_REFERENCE, j->zscr, j->zscr_bytes,
                                                   st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t enqueue_round(apirl_batch* b) {
  BatchHeader* hdr = (BatchHeader*)b->d_tab;
  hipError_t e;
  if ((e = launch_group_clear(hdr, (const ClearJob*)(b->d_tab + b->off_clear), b->grid,
                              b->stream)) != hipSuccess)
    return e;
  if ((e = launch_group_project(hdr, (const ProjJob*)(b->d_tab + b->off_p2),
                                (const ProjJob*)(b->d_tab + b->off_p3), b->grid, b->stream)) !=
      hipSuccess)
    return e;
  if ((e = launch_group_gather(hdr, (const ZsJob*)(b->d_tab + b->off_zs),
                               (const GJob*)(b->d_tab + b->off_g2),
                               (const GJob*)(b->d_tab + b->off_g3), b->grid, b->max_nsino,
                               b->stream)) != hipSuccess)
    return e;
  return launch_group_update(hdr, (const UpdJob*)(b->d_tab + b->off_upd), b->grid, b->stream);
}

hipError_t capture(apirl_batch* b) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  hipError_t e1 = enqueue_round(b);
  e = hipStreamEndCapture(b->stream, &b->graph);
  if (e1 != hipSuccess) return e1;
  if (e != hipSuccess) return e;
  return hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
}

}  // namespace

extern "C" {

apirl_status apirl_batch_create(apirl_ctx* ctx, int max_jobs, int max_sinograms,
                                apirl_batch** out) {
  if (!ctx || !out || max_jobs <= 0 || max_sinograms <= 0) return APIRL_ERR_INVALID;
  *out = nullptr;
  (void)hipSetDevice(ctx->device);
  apirl_batch* b = new apirl_batch();
  b->ctx = ctx;
  b->cap = max_jobs;
  b->max_nsino = max_sinograms;
  size_t off = align_up<char>(sizeof(BatchHeader));
  b->off_clear = off;
  off += align_up<char>(sizeof(ClearJob) * 2 * max_jobs);
  b->off_p2 = off;
  off += align_up<char>(sizeof(ProjJob) * max_jobs);
  b->off_p3 = off;
  off += align_up<char>(sizeof(ProjJob) * max_jobs);
  b->off_zs = off;
  off += align_up<char>(sizeof(ZsJob) * max_jobs);
  b->off_g2 = off;
  off += align_up<char>(sizeof(GJob) * max_jobs);
  b->off_g3 = off;
  off += align_up<char>(sizeof(GJob) * max_jobs);
  b->off_upd = off;
  off += align_up<char>(sizeof(UpdJob) * max_jobs);
  b->tab_bytes = off;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&b->d_tab, off) != hipSuccess ||
      hipHostMalloc(&b->h_tab[0], off, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&b->h_tab[1], off, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&b->copy_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->copy_ev[1], hipEventDisableTiming) != hipSuccess) {
    apirl_batch_destroy(b);
    return APIRL_ERR_NOMEM;
  }
  // the first round's table copy is the first work on the stream: nothing to wait for
  (void)hipEventRecord(b->copy_ev[0], b->stream);
  (void)hipEventRecord(b->copy_ev[1], b->stream);
  *out = b;
  return APIRL_OK;
}

apirl_status apirl_batch_destroy(apirl_batch* b) {
  if (!b) return APIRL_ERR_INVALID;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (BJob* j : b->jobs) {
    if (!j) continue;
    release_bufs(b, j);
    if (j->done_ev) (void)hipEventDestroy(j->done_ev);
    delete j;
  }
  for (auto& pr : b->pool) (void)hipFree(pr.second);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_tab) (void)hipFree(b->d_tab);
  for (int i = 0; i < 2; i++) {
    if (b->h_tab[i]) (void)hipHostFree(b->h_tab[i]);
    if (b->copy_ev[i]) (void)hipEventDestroy(b->copy_ev[i]);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return APIRL_OK;
}

const char* apirl_batch_last_error(const apirl_batch* b) { return b ? b->err.c_str() : ""; }

int apirl_batch_job_backprojector(const apirl_plan* p) {
  if (!p) return -1;
  const bool large = (p->geo.dims == 3 && apirl_plan_image_voxels(p) >= kLargeVoxels) ||
                     p->geo.ns * p->geo.nas > 1;
  if (large && p->geo.dims == 3 && p->geo.ns * p->geo.nas == 1 && !apirl::tile_misfit(p))
    return APIRL_BP_TILE;
  return APIRL_BP_GATHER;
}

size_t apirl_batch_job_bytes(const apirl_plan* p, int subsets) {
  if (!p || subsets < 1) return 0;
  const size_t nvox = apirl_plan_image_voxels(p);
  size_t bytes = (size_t)subsets * nvox * sizeof(float) + nvox * sizeof(float) +
                 max_subset_bins(p, subsets) * sizeof(float);
  if (p->geo.dims == 3) {
    // oversampled bins: the largest gather scratch of the plans of one LOR per bin
    // once they are built (the job's submission builds them), the plan's own before
    size_t z = max_gather_scratch(p, subsets);
    for (const apirl_plan* h : p->os_parts) z = std::max(z, max_gather_scratch(h, subsets));
    bytes += z;
  }
  return bytes;
}

apirl_status apirl_batch_set_launch(apirl_batch* b, int grid, int use_graph) {
  if (!b || grid < 0) return APIRL_ERR_INVALID;
  if (grid > 0 && grid != b->grid) {
    b->grid = grid;
    if (b->exec) {
      (void)hipStreamSynchronize(b->stream);
      (void)hipGraphExecDestroy(b->exec);
      (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
    }
  }
  b->use_graph = use_graph != 0;
  return APIRL_OK;
}

apirl_status apirl_batch_set_schedule(apirl_batch* b, int grab, int order) {
  if (!b) return APIRL_ERR_INVALID;
  if (grab < 1 || grab > 64) return bfail(b, APIRL_ERR_INVALID, "grab must be in 1..64");
  if (order != APIRL_BATCH_ORDER_SUBMISSION && order != APIRL_BATCH_ORDER_LARGEST_FIRST)
    return bfail(b, APIRL_ERR_INVALID, "unknown job order");
  // both are read from the round's table, so a captured round needs no new capture
  b->grab = grab;
  b->order_mode = order;
  return APIRL_OK;
}

apirl_status apirl_batch_submit(apirl_batch* b, const apirl_job_desc* d, int* job_id) {
  if (!b || !d || !job_id || !d->plan || !d->d_meas || !d->d_init || !d->d_out ||
      d->iterations < 1 || d->subsets < 1)
    return APIRL_ERR_INVALID;
  apirl_plan* p = d->plan;
  if (p->ctx != b->ctx) return bfail(b, APIRL_ERR_INVALID, "plan belongs to another context");
  if (d->subsets > p->geo.nang) return bfail(b, APIRL_ERR_INVALID, "more subsets than angles");
  // Oversampled bins (ns x nas LORs): the job runs with the single-job kernels, its
  // backprojection the gathers of the plans of one LOR per bin (apirl_plan::os_parts).
  const bool os = p->geo.ns * p->geo.nas > 1;
  if (os && !apirl::ensure_os_parts(p))
    return bfail(b, APIRL_ERR_UNSUPPORTED, std::string("oversampled job: ") + p->ctx->err);
  std::vector<apirl_plan*> gplans = os ? p->os_parts : std::vector<apirl_plan*>{p};
  for (apirl_plan* q : gplans) {
    if (!q->gather_ok)
      return bfail(b, APIRL_ERR_UNSUPPORTED, "batched jobs use the gather backprojector");
    if (const char* why = apirl::gather_misfit(q, 0, d->subsets))  // subset 0: most angles
      return bfail(b, APIRL_ERR_UNSUPPORTED, std::string("gather: ") + why);
  }
  if (p->geo.dims == 3 && p->geo.nsino > b->max_nsino)
    return bfail(b, APIRL_ERR_INVALID, "sinogram count above the batch's max_sinograms");
  if ((int)b->order.size() >= b->cap) return bfail(b, APIRL_ERR_INVALID, "batch is full");
  (void)hipSetDevice(b->ctx->device);
  BJob* j = new BJob();
  j->plan = p;
  j->T = d->iterations;
  j->S = d->subsets;
  j->osem = d->osem_rules ? 1 : 0;
  j->is2d = p->geo.dims == 2;
  j->meas = d->d_meas;
  j->add = d->d_additive;
  j->add_local = d->additive_own_angles ? 0 : 1;
  // OSEM on the mMR traces the reference's subset LORs (apirl_em_reconstruct alike)
  j->form = (p->scanner == APIRL_SCANNER_MMR && j->osem) ? APIRL_SCANNER_MMR_SUBSET : p->scanner;
  if (!apirl::build_lor_form(p, j->form)) {
    delete j;
    return bfail(b, APIRL_ERR_NOMEM, std::string("job set-up failed: ") + p->ctx->err);
  }
  if (!os && !p->forms[j->form].gather) {
    delete j;
    return bfail(b, APIRL_ERR_UNSUPPORTED, "batched jobs use the gather backprojector");
  }
  j->x = d->d_out;
  j->nvox = apirl_plan_image_voxels(p);
  j->strict = (!j->osem && j->is2d) ? 1 : 0;
  const size_t sens_b = (size_t)j->S * j->nvox * sizeof(float);
  const size_t ratio_b = max_subset_bins(p, j->S) * sizeof(float);
  const size_t bp_b = j->nvox * sizeof(float);
  size_t z_b = 0;
  if (!j->is2d)
    for (apirl_plan* q : gplans) z_b = std::max(z_b, max_gather_scratch(q, j->S));
  auto get = [&](size_t n) -> void* {
    void* q = balloc(b, n);
    if (q) j->bufs.emplace_back(q, n);
    return q;
  };
  j->sens = (float*)get(sens_b);
  j->ratio = (float*)get(ratio_b);
  j->bp = (float*)get(bp_b);
  if (z_b) j->zscr = get(z_b);
  j->zscr_bytes = z_b;
  j->large = (!j->is2d && j->nvox >= kLargeVoxels) || os;
  {
    apirl::ActiveForm fs(p, j->form);
    j->tile = apirl_batch_job_backprojector(p) == APIRL_BP_TILE;
  }
  if (!j->sens || !j->ratio || !j->bp || (z_b && !j->zscr) ||
      hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming) != hipSuccess) {
    release_bufs(b, j);
    delete j;
    return bfail(b, APIRL_ERR_NOMEM, "cannot allocate the job's device state");
  }
  // Work of the caller's stream (inputs) comes first; the job's set-up then runs on
  // the batch stream: sensitivity per subset (OsemSinogram3d.cpp:101-117,
  // MlemSinogram3d.cpp:312-331), MLEM thresholds, and the initial image.
  apirl_ctx* c = b->ctx;
  hipEvent_t ready;
  if (hipEventCreateWithFlags(&ready, hipEventDisableTiming) != hipSuccess) {
    release_bufs(b, j);
    delete j;
    return bfail(b, APIRL_ERR_HIP, "event creation failed");
  }
  (void)hipEventRecord(ready, c->stream);
  (void)hipStreamWaitEvent(b->stream, ready, 0);
  (void)hipEventDestroy(ready);
  hipStream_t saved = c->stream;
  c->stream = b->stream;
  apirl_status st = APIRL_OK;
  j->thr.assign(j->S, 0.f);
  apirl::ActiveForm form_scope(p, j->form);
  for (int s = 0; s < j->S && st == APIRL_OK; s++) {
    st = apirl_sensitivity(p, d->d_mult, j->sens + (size_t)s * j->nvox, s, j->S,
                           j->tile ? APIRL_BP_TILE : APIRL_BP_GATHER, j->osem);
    if (st == APIRL_OK && !j->osem)
      st = apirl_mlem_threshold(c, j->sens + (size_t)s * j->nvox, j->nvox, &j->thr[s]);
  }
  if (st == APIRL_OK &&
      hipMemcpyAsync(j->x, d->d_init, j->nvox * sizeof(float), hipMemcpyDeviceToDevice,
                     b->stream) != hipSuccess)
    st = APIRL_ERR_HIP;
  c->stream = saved;
  if (st != APIRL_OK) {
    release_bufs(b, j);
    (void)hipEventDestroy(j->done_ev);
    delete j;
    return bfail(b, st, std::string("job set-up failed: ") + c->err);
  }
  j->id = (int)b->jobs.size();
  j->active = true;
  b->jobs.push_back(j);
  b->order.push_back(j->id);
  *job_id = j->id;
  return APIRL_OK;
}

apirl_status apirl_batch_step(apirl_batch* b, int* active) {
  if (!b) return APIRL_ERR_INVALID;
  (void)hipSetDevice(b->ctx->device);
  if (b->order.empty()) {
    if (active) *active = 0;
    return APIRL_OK;
  }
  char* h = b->h_tab[b->cur];
  if (hipEventSynchronize(b->copy_ev[b->cur]) != hipSuccess)
    return bfail(b, APIRL_ERR_HIP, "table staging wait failed");
  BatchHeader* hdr = (BatchHeader*)h;
  std::memset(hdr, 0, sizeof(BatchHeader));
  hdr->grab = b->grab;
  ClearJob* cl = (ClearJob*)(h + b->off_clear);
  ProjJob* p2 = (ProjJob*)(h + b->off_p2);
  ProjJob* p3 = (ProjJob*)(h + b->off_p3);
  ZsJob* zs = (ZsJob*)(h + b->off_zs);
  GJob* g2 = (GJob*)(h + b->off_g2);
  GJob* g3 = (GJob*)(h + b->off_g3);
  UpdJob* up = (UpdJob*)(h + b->off_upd);
  std::vector<BJob*> finishing;
  std::vector<std::pair<BJob*, int>> direct;  // large jobs and their subset this round
  // The jobs' virtual blocks are laid out, and so drawn, in this order: submission,
  // or the round's largest sub-iteration first (rays traced plus voxels) so that the
  // small jobs fill the end of each phase.  Results do not depend on it.
  std::vector<int> round_order(b->order);
  if (b->order_mode == APIRL_BATCH_ORDER_LARGEST_FIRST) {
    auto work = [&](int id) {
      const BJob* j = b->jobs[id];
      const Geo& g = j->plan->geo;
      const double nrays = (double)g.nsino * subset_num_angles(g.nang, j->S, j->s) * g.nr;
      return nrays * (g.nx + g.ny + (j->is2d ? 0 : g.nz)) + (double)j->nvox;
    };
    std::stable_sort(round_order.begin(), round_order.end(),
                     [&](int x, int y) { return work(x) > work(y); });
  }
  for (int id : round_order) {
    BJob* j = b->jobs[id];
    apirl::ActiveForm form_scope(j->plan, j->form);
    const apirl_plan* p = j->plan;
    const Geo& g = p->geo;
    const int s = j->s, S = j->S;
    if (j->large) {
      direct.emplace_back(j, s);
      if (++j->s == S) {
        j->s = 0;
        if (++j->t == j->T) finishing.push_back(j);
      }
      continue;
    }
    const int nang_sub = subset_num_angles(g.nang, S, s);
    const long nrays = (long)g.nsino * nang_sub * g.nr;
    // clear the backprojection (and the gather's non-zero flags)
    {
      ClearJob& c = cl[hdr->count[PH_CLEAR]++];
      c.p = j->bp;
      c.n = (long)j->nvox;
      c.first = hdr->total[PH_CLEAR];
      hdr->total[PH_CLEAR] += group_blocks(c.n);
    }
    GatherArgs A;
    ZRay* zr = nullptr;
    int* nzp = nullptr;
    gather_args_for(p, s, S, BP_PLAIN, j->ratio, j->bp, j->zscr, A, zr, nzp);
    if (!j->is2d) {
      ClearJob& c = cl[hdr->count[PH_CLEAR]++];
      c.p = (float*)nzp;
      c.n = (long)nang_sub * g.nr;
      c.first = hdr->total[PH_CLEAR];
      hdr->total[PH_CLEAR] += group_blocks(c.n);
    }
    // fused projection + ratio (Sinogram3D.cpp:830-856, 857+)
    {
      const int ph = j->is2d ? PH_PROJ2 : PH_PROJ3;
      ProjJob& q = (j->is2d ? p2 : p3)[hdr->count[ph]++];
      q.g = g;
      q.rays = p->d_rays;
      q.zab = p->d_zab;
      q.zorder = j->is2d ? nullptr : p->d_zorder;
      q.img = j->x;
      q.fr = FusedRatio();
      q.fr.meas = j->meas;
      q.fr.additive = j->add;
      q.fr.additive_local = j->add_local;
      q.fr.ratio = j->ratio;
      q.fr.lik_partials = nullptr;
      q.fr.lik_positive_only = j->is2d ? 1 : 0;
      q.nang_sub = nang_sub;
      q.subset = s;
      q.nsubsets = S;
      q.pad = 0;
      q.nrays = nrays;
      q.first = hdr->total[ph];
      hdr->total[ph] += group_blocks(nrays);
    }
    if (!j->is2d) {
      ZsJob& z = zs[hdr->count[PH_ZSETUP]++];
      z.g = g;
      z.rays = p->d_rays;
      z.zab = p->d_zab;
      z.sino = j->ratio;
      z.out = zr;
      z.nz_pair = nzp;
      z.ztab = A.ztab;
      z.zsoff = A.zsoff;
      z.zblock = A.zblock;
      z.pad = 0;
      z.kind = BP_PLAIN;
      z.nang_sub = nang_sub;
      z.subset = s;
      z.nsubsets = S;
      z.nrays = nrays;
      z.first = hdr->total[PH_ZSETUP];
      hdr->total[PH_ZSETUP] += group_blocks(nrays);
    }
    {
      const int ph = j->is2d ? PH_GATHER2 : PH_GATHER3;
      GJob& q = (j->is2d ? g2 : g3)[hdr->count[ph]++];
      q.A = A;
      q.ntx = (g.nx + 15) / 16;
      q.nty = (g.ny + 15) / 16;
      q.first = hdr->total[ph];
      hdr->total[ph] += gather_tile_blocks(g);
    }
    {
      UpdJob& u = up[hdr->count[PH_UPDATE]++];
      u.x = j->x;
      u.b = j->bp;
      u.s = j->sens + (size_t)s * j->nvox;
      u.n = (long)j->nvox;
      u.thr = j->thr[s];
      u.strict = j->strict;
      u.first = hdr->total[PH_UPDATE];
      hdr->total[PH_UPDATE] += group_blocks(u.n);
    }
    // advance the job
    if (++j->s == S) {
      j->s = 0;
      if (++j->t == j->T) finishing.push_back(j);
    }
  }
  hipError_t e = hipSuccess;
  if (direct.size() < b->order.size()) {  // some jobs run in the grouped launches
    if (hipMemcpyAsync(b->d_tab, h, b->tab_bytes, hipMemcpyHostToDevice, b->stream) !=
        hipSuccess)
      return bfail(b, APIRL_ERR_HIP, "table upload failed");
    (void)hipEventRecord(b->copy_ev[b->cur], b->stream);
    b->cur ^= 1;
    if (b->use_graph) {
      if (!b->exec && (e = capture(b)) != hipSuccess)
        return bfail(b, APIRL_ERR_HIP,
                     std::string("graph capture failed: ") + hipGetErrorString(e));
      e = hipGraphLaunch(b->exec, b->stream);
    } else {
      e = enqueue_round(b);
    }
    if (e != hipSuccess)
      return bfail(b, APIRL_ERR_HIP, std::string("round launch failed: ") + hipGetErrorString(e));
  }
  // large jobs: the sub-iteration apirl_em_reconstruct runs for them (fused projection
  // and ratio, gather backprojection, update), on the batch stream
  for (auto& pr : direct) {
    BJob* j = pr.first;
    const int s = pr.second;
    apirl::ActiveForm form_scope(j->plan, j->form);
    FusedRatio fr;
    fr.meas = j->meas;
    fr.additive = j->add;
    fr.additive_local = j->add_local;
    fr.ratio = j->ratio;
    fr.lik_partials = nullptr;
    fr.lik_positive_only = 0;
    if ((e = launch_project(j->plan, j->x, nullptr, s, j->S, APIRL_PREC_REFERENCE, &fr, nullptr,
                            b->stream)) != hipSuccess)
      return bfail(b, APIRL_ERR_HIP, std::string("large-job launch failed: ") +
                                         hipGetErrorString(e));
    if (j->tile) {
      // the tile backprojection with the update fused in, as apirl_em_reconstruct runs
      // it with APIRL_BP_TILE (on the batch stream; the plan keeps its geometry)
      apirl_ctx* c = b->ctx;
      hipStream_t saved = c->stream;
      c->stream = b->stream;
      const apirl_status st = apirl_backproject_update(
          j->plan, j->ratio, j->x, j->sens + (size_t)s * j->nvox, j->thr[s], j->strict, s, j->S,
          APIRL_BP_TILE, APIRL_PREC_REFERENCE);
      c->stream = saved;
      if (st != APIRL_OK)
        return bfail(b, st, std::string("large-job tile backprojection failed: ") + c->err);
      continue;
    }
    if (j->plan->geo.ns * j->plan->geo.nas == 1) {
      // the update fused into the gather, as apirl_em_reconstruct runs it
      const EmFuse fz{j->sens + (size_t)s * j->nvox, j->thr[s], j->strict};
      if ((e = launch_backproject_gather(j->plan, j->ratio, j->x, s, j->S, BP_PLAIN,
                                         APIRL_PREC_REFERENCE, j->zscr, j->zscr_bytes, b->stream,
                                         &fz)) != hipSuccess)
        return bfail(b, APIRL_ERR_HIP, std::string("large-job launch failed: ") +
                                           hipGetErrorString(e));
      continue;
    }
    if ((e = hipMemsetAsync(j->bp, 0, j->nvox * sizeof(float), b->stream)) != hipSuccess ||
        (e = backproject_job(j, s, b->stream)) != hipSuccess ||
        (e = launch_em_update(j->x, j->bp, j->sens + (size_t)s * j->nvox, j->nvox, j->thr[s],
                              j->strict, b->stream)) != hipSuccess)
      return bfail(b, APIRL_ERR_HIP, std::string("large-job launch failed: ") +
                                         hipGetErrorString(e));
  }
  for (BJob* j : finishing) {
    (void)hipEventRecord(j->done_ev, b->stream);
    j->done_recorded = true;
    j->active = false;
    b->order.erase(std::find(b->order.begin(), b->order.end(), j->id));
  }
  if (active) *active = (int)b->order.size();
  return APIRL_OK;
}

apirl_status apirl_batch_job_done(apirl_batch* b, int id, int* done) {
  if (!b || !done || id < 0 || id >= (int)b->jobs.size() || !b->jobs[id])
    return APIRL_ERR_INVALID;
  BJob* j = b->jobs[id];
  *done = 0;
  if (!j->done_recorded) return APIRL_OK;
  const hipError_t e = hipEventQuery(j->done_ev);
  if (e == hipSuccess) *done = 1;
  else if (e != hipErrorNotReady) return bfail(b, APIRL_ERR_HIP, "event query failed");
  return APIRL_OK;
}

apirl_status apirl_batch_synchronize(apirl_batch* b) {
  if (!b) return APIRL_ERR_INVALID;
  return hipStreamSynchronize(b->stream) == hipSuccess ? APIRL_OK : APIRL_ERR_HIP;
}

apirl_status apirl_batch_release(apirl_batch* b, int id) {
  if (!b || id < 0 || id >= (int)b->jobs.size() || !b->jobs[id]) return APIRL_ERR_INVALID;
  BJob* j = b->jobs[id];
  int done = 0;
  apirl_status st = apirl_batch_job_done(b, id, &done);
  if (st != APIRL_OK) return st;
  if (!done) return bfail(b, APIRL_ERR_INVALID, "job still running");
  release_bufs(b, j);
  (void)hipEventDestroy(j->done_ev);
  delete j;
  b->jobs[id] = nullptr;
  return APIRL_OK;
}

}  // extern "C"
