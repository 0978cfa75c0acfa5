// Snapshot batches: many fields of uneven grid sizes reconstructed together.
//
// Each snapshot owns a context (field, work buffers, device scalars with its
// own stop flag); snapshots of one grid share one geometry plan.  A run
// either queues every snapshot's reconstruction into one hipGraph --
// captured once over a few fork-join branches (streams), then replayed --
// or queues them one after the other with a synchronisation each (the
// one-launch-per-snapshot schedule it is compared with).  Convergence is
// per snapshot: a snapshot whose stop rule fired turns its remaining
// iteration launches into no-ops, inside the graph as well.
//
// The autotuner picks, per distinct grid, the tile edge of the tile engine
// (8 or 16 voxels, which also bounds a segment at 22 or 46 visits) and the
// workgroup size of its passes (256, 512, 1024) by timing the candidates on
// the first snapshot of that grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "driver.hpp"
#include "host.hpp"
#include "launch.hpp"

using namespace omni3d;

struct omni3d_batch {
  omni3d_params base{};
  std::vector<omni3d_ctx*> snap;
  std::vector<std::unique_ptr<DevBuf>> out;  // FP64 result per snapshot
  std::vector<int> branch;   // graph branch of each snapshot
  hipStream_t main = nullptr;
  std::vector<hipStream_t> br;
  std::vector<hipEvent_t> br_ev;
  hipEvent_t fork = nullptr, fork2 = nullptr, t0 = nullptr, t1 = nullptr;
  // one launch per source iteration for every snapshot (when all can):
  // the snapshots' descriptors and the snapshot of each workgroup
  DevBuf desc, blk;
  unsigned int nblk = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool prepared = false;     // plans shared, buffers sized
  std::map<std::array<int, 3>, std::array<int, 2>> tuned;  // grid -> (tile bits, threads)
  std::string err;
  ~omni3d_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (omni3d_ctx* c : snap) omni3d_destroy(c);
    for (hipEvent_t e : br_ev) (void)hipEventDestroy(e);
    for (hipStream_t s : br) (void)hipStreamDestroy(s);
    if (fork) (void)hipEventDestroy(fork);
    if (fork2) (void)hipEventDestroy(fork2);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (main) (void)hipStreamDestroy(main);
  }
};

namespace {

constexpr int kBranches = 4;

int bfail(omni3d_batch* b, int code, const std::string& msg) {
  if (b) b->err = msg;
  return fail(nullptr, code, msg);
}

#define BATCH_TRY_HIP(b, expr)                                                       \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess)                                                            \
      return bfail((b), OMNI3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

void drop_graph(omni3d_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

std::array<int, 3> dims_of(const omni3d_ctx* c) { return {c->prm.nx, c->prm.ny, c->prm.nz}; }

// One untimed and one timed reconstruction of snapshot i on the main stream;
// returns the timed milliseconds (or < 0 on error).
double time_snapshot(omni3d_batch* b, int i) {
  omni3d_ctx* c = b->snap[i];
  c->stream = b->main;
  if (build_plan(c)) return -1.0;
  for (int rep = 0; rep < 2; rep++) {
    if (rep == 1 && hipEventRecord(b->t0, b->main) != hipSuccess) return -1.0;
    if (enqueue_solve(c, -1, b->out[i]->as<void>(), nullptr, OMNI3D_F64)) return -1.0;
    if (rep == 1 && hipEventRecord(b->t1, b->main) != hipSuccess) return -1.0;
    if (hipStreamSynchronize(b->main) != hipSuccess) return -1.0;
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, b->t0, b->t1) != hipSuccess) return -1.0;
  return ms;
}

// Shares one plan among the snapshots of each grid, sizes every buffer with
// an eager run, and assigns the graph branches (largest first, to the least
// loaded branch).
int prepare(omni3d_batch* b) {
  if (b->prepared) return OMNI3D_OK;
  std::map<std::array<int, 3>, int> first;
  for (int i = 0; i < (int)b->snap.size(); i++) {
    omni3d_ctx* c = b->snap[i];
    if (!c->have_field) return bfail(b, OMNI3D_ERR_STATE, "snapshot " + std::to_string(i) + " has no field");
    const auto d = dims_of(c);
    auto t = b->tuned.find(d);
    if (t != b->tuned.end()) {
      c->tile_bits = t->second[0];
      c->tile_threads = t->second[1];
    }
    c->stream = b->main;
    auto f = first.find(d);
    if (f == first.end()) {
      first[d] = i;
      const int rc = build_plan(c);
      if (rc) return bfail(b, rc, c->err);
    } else {
      omni3d_ctx* owner = b->snap[f->second];
      c->plan = owner->plan;
      c->dirs_dirty = false;
      c->tile_bits = owner->tile_bits;
      c->tile_threads = owner->tile_threads;
    }
  }
  for (int i = 0; i < (int)b->snap.size(); i++) {
    const int rc = enqueue_solve(b->snap[i], -1, b->out[i]->as<void>(), nullptr, OMNI3D_F64);
    if (rc) return bfail(b, rc, b->snap[i]->err);
  }
  BATCH_TRY_HIP(b, hipStreamSynchronize(b->main));
  std::vector<int> order(b->snap.size());
  for (int i = 0; i < (int)order.size(); i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    return voxels(b->snap[x]->prm) > voxels(b->snap[y]->prm);
  });
  std::vector<int64_t> load(kBranches, 0);
  b->branch.assign(b->snap.size(), 0);
  for (int i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->branch[i] = k;
    load[k] += voxels(b->snap[i]->prm);
  }
  b->prepared = true;
  return OMNI3D_OK;
}

// Uploads the iteration descriptors when every snapshot's source iteration
// can share one launch per iteration (same precision for all); else 0.
int fused_descriptors(omni3d_batch* b, bool* fused) {
  *fused = false;
  if (std::getenv("OMNI3D_BATCH_SEPARATE")) return OMNI3D_OK;
  for (omni3d_ctx* c : b->snap)
    if (!affine_fusable(c) || c->prm.precision != b->snap[0]->prm.precision ||
        c->plan->dense_group != b->snap[0]->plan->dense_group)
      return OMNI3D_OK;
  std::vector<IterDesc> d;
  std::vector<unsigned int> blk;
  for (int i = 0; i < (int)b->snap.size(); i++) {
    d.push_back(affine_iter_desc(b->snap[i], (int)blk.size()));
    blk.insert(blk.end(), (size_t)d.back().nb, (unsigned int)i);
  }
  BATCH_TRY_HIP(b, b->desc.alloc(sizeof(IterDesc) * d.size()));
  BATCH_TRY_HIP(b, b->blk.alloc(sizeof(unsigned int) * blk.size()));
  BATCH_TRY_HIP(b, hipMemcpy(b->desc.as<void>(), d.data(), sizeof(IterDesc) * d.size(),
                             hipMemcpyHostToDevice));
  BATCH_TRY_HIP(b, hipMemcpy(b->blk.as<void>(), blk.data(), sizeof(unsigned int) * blk.size(),
                             hipMemcpyHostToDevice));
  b->nblk = (unsigned int)blk.size();
  *fused = true;
  return OMNI3D_OK;
}

// Queues every snapshot's reconstruction (phase: as in run_affine) on its
// branch, forked from and joined back into the main stream.
int enqueue_branches(omni3d_batch* b, int phase, hipEvent_t fork_ev) {
  hipError_t e = hipEventRecord(fork_ev, b->main);
  for (int k = 0; e == hipSuccess && k < kBranches; k++) e = hipStreamWaitEvent(b->br[k], fork_ev, 0);
  if (e != hipSuccess) return bfail(b, OMNI3D_ERR_HIP, std::string("fork: ") + hipGetErrorString(e));
  int rc = OMNI3D_OK;
  for (int i = 0; rc == OMNI3D_OK && i < (int)b->snap.size(); i++) {
    omni3d_ctx* c = b->snap[i];
    c->stream = b->br[b->branch[i]];
    c->capturing = true;
    c->affine_phase = phase;
    rc = enqueue_solve(c, -1, b->out[i]->as<void>(), nullptr, OMNI3D_F64);
    c->affine_phase = 0;
    c->capturing = false;
    c->stream = b->main;
  }
  if (rc) return bfail(b, rc, "capture: " + g_last_error);
  for (int k = 0; e == hipSuccess && k < kBranches; k++) {
    e = hipEventRecord(b->br_ev[k], b->br[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->main, b->br_ev[k], 0);
  }
  if (e != hipSuccess) return bfail(b, OMNI3D_ERR_HIP, std::string("join: ") + hipGetErrorString(e));
  return OMNI3D_OK;
}

// Captures the batch: passes A of all snapshots over the branches, then --
// when all snapshots allow it -- one launch per source iteration for every
// snapshot on the main stream, then passes B over the branches; otherwise
// each branch runs its snapshots' whole reconstructions.
int capture(omni3d_batch* b) {
  drop_graph(b);
  bool fused = false;
  int rc = fused_descriptors(b, &fused);
  if (rc) return rc;
  BATCH_TRY_HIP(b, hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
  rc = enqueue_branches(b, fused ? 1 : 0, b->fork);
  if (rc == OMNI3D_OK && fused) {
    const omni3d_params& P = b->snap[0]->prm;
    const int inc = P.precision == OMNI3D_FP32_KAHAN && !std::getenv("OMNI3D_NO_INCREMENTS");
    hipError_t e = hipSuccess;
    for (int k = 1; e == hipSuccess && k <= P.max_iterations; k++)
      e = launch_iterate_multi(b->snap[0]->plan->dense_group, inc, k, b->desc.as<IterDesc>(),
                               b->blk.as<unsigned int>(), b->nblk,
                               P.max_iterations, P.tolerance, b->main);
    if (e != hipSuccess) rc = bfail(b, OMNI3D_ERR_HIP, std::string("iterations: ") + hipGetErrorString(e));
    if (rc == OMNI3D_OK) rc = enqueue_branches(b, 2, b->fork2);
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->main, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bfail(b, OMNI3D_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e2));
  }
  b->graph = g;
  BATCH_TRY_HIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return OMNI3D_OK;
}

}  // namespace

extern "C" {

int omni3d_batch_create(const omni3d_params* base, int count, const int* nx, const int* ny,
                        const int* nz, omni3d_batch** out) {
  if (!base || !out || !nx || !ny || !nz || count < 1)
    return fail(nullptr, OMNI3D_ERR_ARG, "batch: null argument or empty batch");
  omni3d_batch* b = new omni3d_batch();
  b->base = *base;
  for (int i = 0; i < count; i++) {
    omni3d_params p = *base;
    p.nx = nx[i];
    p.ny = ny[i];
    p.nz = nz[i];
    omni3d_ctx* c = nullptr;
    const int rc = omni3d_create(&p, &c);
    if (rc) {
      delete b;
      return rc;
    }
    // the masked methods' affine schedule sizes its rows from the field on
    // the host (a synchronisation per field): a captured graph replays the
    // reference schedule instead
    c->affine_masked = false;
    // the snapshots' iterations share one launch: no group count rounded up
    // to fill the CUs per snapshot
    c->groups_fill_cus = false;
    b->snap.push_back(c);
  }
  for (int i = 0; i < count; i++) b->out.emplace_back(new DevBuf());
  hipError_t e = hipStreamCreate(&b->main);
  for (int k = 0; e == hipSuccess && k < kBranches; k++) {
    hipStream_t s;
    hipEvent_t ev;
    e = hipStreamCreate(&s);
    if (e == hipSuccess) b->br.push_back(s);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) b->br_ev.push_back(ev);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork2, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->t0);
  if (e == hipSuccess) e = hipEventCreate(&b->t1);
  for (int i = 0; e == hipSuccess && i < count; i++)
    e = b->out[i]->alloc(sizeof(double) * voxels(b->snap[i]->prm));
  if (e != hipSuccess) {
    delete b;
    return fail(nullptr, OMNI3D_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
  }
  *out = b;
  return OMNI3D_OK;
}

void omni3d_batch_destroy(omni3d_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->main);
  delete b;
}

const char* omni3d_batch_last_error(const omni3d_batch* b) {
  return b ? b->err.c_str() : omni3d_last_global_error();
}

int omni3d_batch_set_directions(omni3d_batch* b, const float* k1, const float* k2,
                                const float* k3, int count) {
  if (!b) return fail(nullptr, OMNI3D_ERR_ARG, "null batch");
  for (omni3d_ctx* c : b->snap) {
    const int rc = omni3d_set_directions(c, k1, k2, k3, count);
    if (rc) return bfail(b, rc, c->err);
  }
  drop_graph(b);
  b->prepared = false;
  return OMNI3D_OK;
}

int omni3d_batch_set_field(omni3d_batch* b, int index, const void* dudt, const void* dvdt,
                           const void* dwdt, const void* mask, int dtype, int mem_kind) {
  if (!b || index < 0 || index >= (int)b->snap.size())
    return fail(nullptr, OMNI3D_ERR_ARG, "batch: bad snapshot index");
  omni3d_ctx* c = b->snap[index];
  c->stream = b->main;
  const void* old = c->field.as<void>();
  const void* old_mask = c->mask.as<void>();
  const bool had_mask = c->have_mask;
  const int rc = omni3d_set_field(c, dudt, dvdt, dwdt, mask, dtype, mem_kind);
  if (rc) return bfail(b, rc, c->err);
  // the graph holds buffer addresses and was captured with or without the
  // mask: a reallocated field or mask, or a mask added or removed,
  // invalidates it
  if (c->field.as<void>() != old || c->have_mask != had_mask ||
      (c->have_mask && c->mask.as<void>() != old_mask))
    drop_graph(b);
  // a mask added or removed changes the work buffers the masked methods
  // use: size them again with an eager run before the next capture
  if (c->have_mask != had_mask) b->prepared = false;
  return OMNI3D_OK;
}

int omni3d_batch_tune(omni3d_batch* b, int* tile_bits_out, int* threads_out) {
  if (!b) return fail(nullptr, OMNI3D_ERR_ARG, "null batch");
  std::map<std::array<int, 3>, int> first;
  for (int i = 0; i < (int)b->snap.size(); i++) {
    if (!b->snap[i]->have_field) return bfail(b, OMNI3D_ERR_STATE, "snapshot without field");
    first.emplace(dims_of(b->snap[i]), i);
  }
  int n = 0;
  for (const auto& kv : first) {
    omni3d_ctx* c = b->snap[kv.second];
    double best = 1e300;
    std::array<int, 2> pick = {c->tile_bits, c->tile_threads};
    for (int bits = kTileBitsMin; bits <= kTileBitsMax; bits++) {
      for (int threads : {256, 512, 1024}) {
        c->tile_bits = bits;
        c->tile_threads = threads;
        const double ms = time_snapshot(b, kv.second);
        if (ms < 0) return bfail(b, OMNI3D_ERR_HIP, "tuning run failed: " + c->err);
        if (ms < best) {
          best = ms;
          pick = {bits, threads};
        }
      }
    }
    b->tuned[kv.first] = pick;
    c->tile_bits = pick[0];
    c->tile_threads = pick[1];
    if (tile_bits_out) tile_bits_out[n] = pick[0];
    if (threads_out) threads_out[n] = pick[1];
    n++;
  }
  // plans are rebuilt with the chosen tiles and shared again
  for (omni3d_ctx* c : b->snap) {
    if (c->plan.use_count() > 1) c->plan = std::make_shared<Plan>();
    c->plan->valid = false;
  }
  drop_graph(b);
  b->prepared = false;
  return n;
}

int omni3d_batch_run(omni3d_batch* b, int use_graph, double* ms_out) {
  if (!b) return fail(nullptr, OMNI3D_ERR_ARG, "null batch");
  int rc = prepare(b);
  if (rc) return rc;
  if (use_graph && !b->exec) {
    rc = capture(b);
    if (rc) return rc;
  }
  BATCH_TRY_HIP(b, hipEventRecord(b->t0, b->main));
  if (use_graph) {
    BATCH_TRY_HIP(b, hipGraphLaunch(b->exec, b->main));
  } else {
    // one reconstruction after the other, each waited for (the schedule of a
    // caller that reconstructs snapshot by snapshot)
    for (int i = 0; i < (int)b->snap.size(); i++) {
      omni3d_ctx* c = b->snap[i];
      c->stream = b->main;
      rc = enqueue_solve(c, -1, b->out[i]->as<void>(), nullptr, OMNI3D_F64);
      if (rc) return bfail(b, rc, c->err);
      BATCH_TRY_HIP(b, hipStreamSynchronize(b->main));
    }
  }
  BATCH_TRY_HIP(b, hipEventRecord(b->t1, b->main));
  BATCH_TRY_HIP(b, hipStreamSynchronize(b->main));
  if (ms_out) {
    float ms = 0;
    BATCH_TRY_HIP(b, hipEventElapsedTime(&ms, b->t0, b->t1));
    *ms_out = ms;
  }
  return OMNI3D_OK;
}

int omni3d_batch_result(omni3d_batch* b, int index, void* p, int dtype, int mem_kind,
                        omni3d_report* rep) {
  if (!b || index < 0 || index >= (int)b->snap.size())
    return fail(nullptr, OMNI3D_ERR_ARG, "batch: bad snapshot index");
  omni3d_ctx* c = b->snap[index];
  const int64_t V = voxels(c->prm);
  if (p) {
    if (dtype == OMNI3D_F64) {
      BATCH_TRY_HIP(b, hipMemcpy(p, b->out[index]->as<void>(), sizeof(double) * V,
                                 mem_kind == OMNI3D_MEM_HOST ? hipMemcpyDeviceToHost
                                                             : hipMemcpyDeviceToDevice));
    } else {
      DevBuf tmp;
      BATCH_TRY_HIP(b, tmp.alloc(sizeof(float) * V));
      BATCH_TRY_HIP(b, (launch_convert<double, float>(b->out[index]->as<double>(), tmp.as<float>(), V,
                                                      b->main)));
      BATCH_TRY_HIP(b, hipStreamSynchronize(b->main));
      BATCH_TRY_HIP(b, hipMemcpy(p, tmp.as<void>(), sizeof(float) * V,
                                 mem_kind == OMNI3D_MEM_HOST ? hipMemcpyDeviceToHost
                                                             : hipMemcpyDeviceToDevice));
    }
  }
  if (rep) {
    DevScalars hs;
    BATCH_TRY_HIP(b, hipMemcpy(&hs, c->scal.as<void>(), sizeof(hs), hipMemcpyDeviceToHost));
    *rep = omni3d_report{};
    rep->iterations = hs.kstar;
    rep->rel_change = hs.rel;
    rep->rays = c->plan->nrays;
    rep->steps = c->plan->steps;
    rep->sources = c->plan->nS;
    rep->status = hs.status;
    rep->source_visits = c->plan->nnz;
    rep->segments = c->plan->nsegs;
    rep->sell_cells = c->plan->ncell;
    rep->matrix_words = c->plan->nwords;
  }
  return OMNI3D_OK;
}

}  // extern "C"
