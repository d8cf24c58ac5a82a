// C ABI of the LEE time advance (include/lee.h): handle lifecycle, the
// reference's init sequence, step orchestration for the fused and plain
// paths, and the unit entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "lee_internal.h"

using namespace lee;

namespace {

std::string g_create_err;

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      h->err = std::string("HIP error ") + hipGetErrorString(_e) + " at " #expr;       \
      return LEE_EHIP;                                                                 \
    }                                                                                  \
  } while (0)

int fail(lee_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  else g_create_err = msg;
  return code;
}

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  return hipMalloc((void **)p, sizeof(T) * count);
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void free_block(Block &b) {
  dfree(b.xyz);
  dfree(b.met);
  dfree(b.jac);
  dfree(b.iblank);
  dfree(b.code_fd);
  dfree(b.clean_fd);
  dfree(b.clean_flt);
  dfree(b.code_flt);
  dfree(b.mean);
  dfree(b.mgrad);
  dfree(b.rhobar);
  dfree(b.mc64);
  dfree(b.mc32);
  dfree(b.rho64);
  dfree(b.rho32);
  for (int k = 0; k < 4; k++) {
    dfree(b.Y64[k]);
    dfree(b.Y32[k]);
  }
  dfree(b.dflux);
  dfree(b.dsum);
  dfree(b.func0);
  dfree(b.func);
  for (auto &z : b.zones) {
    dfree(z.sigma64);
    dfree(z.sigma32);
    dfree(z.hg);
  }
  for (auto &f : b.faces) dfree(f.hg);
  dfree(b.src.g);
  for (int k = 0; k < 5; k++) {
    dfree(b.A64[k]);
    dfree(b.A32[k]);
  }
  dfree(b.amu64);
  dfree(b.amu32);
  dfree(b.agrad);
  dfree(b.obs_l);
  dfree(b.obs_w);
  dfree(b.obs_buf);
  dfree(b.ckpt);
  dfree(b.samp);
}

int create_impl(const lee_config *cfgs, int cfg_stride, const lee_block *blocks, int nblocks, const lee_overset *ov,
                bool batch, lee_handle **out);

bool is_fp32(const lee_handle *h) { return h->cfg.precision == LEE_FP32; }
bool is_plain(const lee_handle *h) { return h->cfg.path == LEE_PATH_PLAIN; }

// Buffer that receives a stage's boundary update in the fused path.
int fused_target(const lee_handle *h, int stage) {
  const int last = h->cfg.rk_stages - 1;
  if (stage == last) return 1;
  return stage % 2 == 0 ? 2 : 3;
}

// Records one event of the timing pool on the handle's stream.
hipError_t mark(lee_handle *h) {
  if (h->evused == h->evpool.size()) {
    hipEvent_t e;
    // device-scope fences only: the timed kernels never touch host memory
    hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    if (r != hipSuccess) return r;
    h->evpool.push_back(e);
  }
  return hipEventRecord(h->evpool[h->evused++], h->stream);
}

int enqueue_fused_step(lee_handle *h) {
  const int S = h->cfg.rk_stages, last = S - 1;
  const double tn = h->time_sol, dt = h->dt;
  const bool f32 = is_fp32(h);
  for (int st = 0; st < S; st++) {
    StageScalars sc{};
    sc.h = dt;
    sc.b = h->b_i[st];
    sc.a = h->a_next[st];
    source_factors(h->cfg, tn + h->c_i[st] * dt, sc.s1, sc.s2);
    // boundary data after the stage: t + c_{min(i+1, last)} h (temporal:236-246)
    const double t_bc = tn + h->c_i[std::min(st + 1, last)] * dt;
    double he_bc = 0.0;
    if (h->cfg.harmonic_wave == LEE_WAVE_ENTROPY) {
      sc.he = harmonic_entropy(h->cfg, tn + h->c_i[st] * dt);
      he_bc = harmonic_entropy(h->cfg, t_bc);
    }
    if (h->ktiming) HIPCHK(mark(h));
    if (pair_stage_ok(h))
      HIPCHK(fused_stage_pair(h->stream, h, st, sc));
    else
      for (int ib = 0; ib < h->nblk; ib++)
        HIPCHK(h->cfg.num_dim == 3 ? fused_stage3d(h->stream, h, h->b[ib], st, sc)
                                   : fused_stage(h->stream, h, h->b[ib], st, sc));
    if (h->ktiming) {
      HIPCHK(mark(h));
      h->evclass.push_back(st == 0 ? 0 : (st == last ? 2 : 1));
    }
    const int tgt = fused_target(h, st);
    bool any_bnd = h->ov.nrecv > 0;
    for (int ib = 0; ib < h->nblk; ib++) any_bnd |= !h->b[ib].faces.empty();
    if (h->ktiming && any_bnd) HIPCHK(mark(h));
    if (any_bnd) HIPCHK(boundary_phase_fused(h->stream, h, f32, tgt, st == last, he_bc));
    if (h->ktiming && any_bnd) {
      HIPCHK(mark(h));
      h->evclass.push_back(4);
    }
  }
  const bool pair_filter = pair_filter_ok(h);
  if (pair_filter) {
    if (h->ktiming) HIPCHK(mark(h));
    HIPCHK(fused_filter_pair(h->stream, h, 1, 0));
    if (h->ktiming) {
      HIPCHK(mark(h));
      h->evclass.push_back(3);
    }
  }
  for (int ib = 0; ib < h->nblk && !pair_filter; ib++) {
    Block &b = h->b[ib];
    if (h->cfg.filter_order > 0) {
      if (h->ktiming) HIPCHK(mark(h));
      HIPCHK(h->cfg.num_dim == 3 ? fused_filter3d(h->stream, h, b, 1, 0, 2) : fused_filter(h->stream, h, b, 1, 0));
      if (h->ktiming) {
        HIPCHK(mark(h));
        h->evclass.push_back(3);
      }
    } else {
      // no filter: y_{n+1} becomes the state (holes are already at sol_ref)
      std::swap(b.Y64[0], b.Y64[1]);
      std::swap(b.Y32[0], b.Y32[1]);
    }
  }
  return LEE_OK;
}

// temporal::time_advance with one kernel per reference operator (FP64).
int enqueue_plain_step(lee_handle *h) {
  const int S = h->cfg.rk_stages, last = S - 1;
  const double tn = h->time_sol, dt = h->dt;
  hipStream_t s = h->stream;
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    const size_t bytes = sizeof(double) * nvars(h) * (size_t)b.n;
    HIPCHK(hipMemcpyAsync(b.Y64[1], b.Y64[0], bytes, hipMemcpyDeviceToDevice, s));  // ynp1 = yn
    HIPCHK(hipMemcpyAsync(b.Y64[2], b.Y64[0], bytes, hipMemcpyDeviceToDevice, s));  // interim = yn
  }
  const bool harm = h->cfg.harmonic_wave == LEE_WAVE_ENTROPY;
  for (int st = 0; st < S; st++) {
    double s1, s2;
    source_factors(h->cfg, tn + h->c_i[st] * dt, s1, s2);
    const double he = harm ? harmonic_entropy(h->cfg, tn + h->c_i[st] * dt) : 0.0;
    const double he_bc = harm ? harmonic_entropy(h->cfg, tn + h->c_i[std::min(st + 1, last)] * dt) : 0.0;
    for (int ib = 0; ib < h->nblk; ib++) {
      Block &b = h->b[ib];
      HIPCHK(plain_rhs(s, h, b, b.Y64[2], b.Y64[3], s1, s2, b.rho64, he));
    }
    for (int ib = 0; ib < h->nblk; ib++) {
      Block &b = h->b[ib];
      HIPCHK(plain_rk_update(s, b, b.Y64[3], b.Y64[1], b.Y64[2], b.Y64[0], dt, h->b_i[st], h->a_next[st],
                             st == last, nvars(h)));
    }
    const int tgt = st < last ? 2 : 1;
    for (int ib = 0; ib < h->nblk; ib++) {
      Block &b = h->b[ib];
      if (!b.faces.empty()) HIPCHK(apply_bcs(s, h, b, b.Y64[tgt], nullptr, false, nullptr, nullptr, he_bc));
    }
    HIPCHK(overset_interpolate(s, h, false, tgt, false));
    for (int ib = 0; ib < h->nblk; ib++) {
      Block &b = h->b[ib];
      HIPCHK(aux_density(s, b, h->cfg.gamma, b.Y64[tgt], b.rho64));
    }
  }
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    if (h->cfg.filter_order > 0) HIPCHK(plain_filter(s, h, b, b.Y64[1]));
    HIPCHK(plain_iblank_reset(s, b, b.Y64[1], nvars(h)));
    std::swap(b.Y64[0], b.Y64[1]);
  }
  return LEE_OK;
}

int enqueue_step(lee_handle *h) { return is_plain(h) ? enqueue_plain_step(h) : enqueue_fused_step(h); }

// One step of a batched ensemble: a persistent launch per RK stage over the
// tiles of every solve, the boundary conditions of the solves that have
// any, then the filter launch, which also advances the device step counter.
// The launches only read operand tables and the schedule, so the sequence
// can be captured once in a HIP graph and replayed.
int enqueue_batch_step(lee_handle *h) {
  const int S = h->cfg.rk_stages, last = S - 1;
  const bool f32 = is_fp32(h);
  for (int st = 0; st < S; st++) {
    if (h->ktiming) HIPCHK(mark(h));
    HIPCHK(batch_stage(h->stream, h, st));
    if (h->ktiming) {
      HIPCHK(mark(h));
      h->evclass.push_back(st == 0 ? 0 : (st == last ? 2 : 1));
    }
    const int tgt = fused_target(h, st);
    for (int ib = 0; ib < h->nblk; ib++) {
      Block &b = h->b[ib];
      if (b.faces.empty()) continue;
      const bool rho = st == last && !b.uniform_mean;
      HIPCHK(apply_bcs(h->stream, h, b, b.Y64[tgt], b.Y32[tgt], f32, rho ? b.rho64 : nullptr,
                       rho ? b.rho32 : nullptr));
    }
  }
  if (h->ktiming) HIPCHK(mark(h));
  HIPCHK(batch_filter(h->stream, h));
  if (h->ktiming) {
    HIPCHK(mark(h));
    h->evclass.push_back(3);
  }
  return LEE_OK;
}

// solver::solve's step prologue for each solve of a batch, `nsteps` steps
// ahead: the clock never depends on the solution (the CFL maximum is a
// function of metrics and mean flow), so dt, time and the source factors
// of every stage are known in advance.
void batch_schedule(lee_handle *h, int nsteps, BatchSched *out) {
  const int S = h->cfg.rk_stages, ns = h->nblk;
  for (int k = 0; k < ns; k++) {
    Block &b = h->b[k];
    for (int st = 0; st < nsteps; st++) {
      b.clk_step++;
      const double cmax = cfl_max_block(b, b.clk_dt, -1e99);
      double dt = b.clk_dt;
      if (b.cfg.fix_cfl) {
        const double ratio = b.cfg.cfl_or_dt / cmax;
        dt *= ratio;
        b.clk_dt = dt;
      }
      b.clk_time += dt;
      for (int i = 0; i < S; i++) {
        BatchSched &q = out[((size_t)st * ns + k) * S + i];
        q.dt = dt;
        source_factors(b.cfg, b.clk_time + h->c_i[i] * dt, q.s1, q.s2);
      }
    }
  }
}

// Makes room for `nsteps` steps of schedule on the device.
int batch_reserve(lee_handle *h, size_t nsteps) {
  BatchPlan &p = h->bp;
  if (p.sched_steps >= nsteps) return LEE_OK;
  // the graph holds the old address
  if (p.exec) HIPCHK(hipGraphExecDestroy(p.exec));
  p.exec = nullptr;
  dfree(p.sched);
  HIPCHK(dalloc(&p.sched, nsteps * h->nblk * h->cfg.rk_stages));
  p.sched_steps = nsteps;
  return LEE_OK;
}

// Captures one batched step in a HIP graph (once per table layout).
int batch_capture(lee_handle *h) {
  BatchPlan &p = h->bp;
  if (p.exec) return LEE_OK;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_batch_step(h);
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIPCHK(e);
  const hipError_t ei = hipGraphInstantiate(&p.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(ei);
  return LEE_OK;
}

// Tile shape and persistent grid of the batched kernels, timed with HIP
// events on scratch tables (the state in Y0 is untouched).
int batch_autotune(lee_handle *h) {
  BatchPlan &p = h->bp;
  if (p.tuned) return LEE_OK;
  int rc = batch_reserve(h, 1);
  if (rc) return rc;
  std::vector<BatchSched> zero(h->nblk * h->cfg.rk_stages, BatchSched{0.0, 0.0, 0.0});
  HIPCHK(hipMemcpy(p.sched, zero.data(), sizeof(BatchSched) * zero.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.counter, h->bp_zero, sizeof(h->bp_zero), hipMemcpyHostToDevice));
  int dev_cu = 256;
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess && v > 0) dev_cu = v;
  }
  const int grids[5] = {0, LEE_GRID_RESIDENT, 4 * dev_cu, 8 * dev_cu, 16 * dev_cu};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int reps = 5;
  auto time_it = [&](bool stage, float &ms) -> int {
    for (int k = 0; k < 2; k++) HIPCHK(stage ? batch_stage(h->stream, h, 1) : batch_filter(h->stream, h));
    HIPCHK(hipEventRecord(e0, h->stream));
    for (int k = 0; k < reps; k++) HIPCHK(stage ? batch_stage(h->stream, h, 1) : batch_filter(h->stream, h));
    HIPCHK(hipEventRecord(e1, h->stream));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    return LEE_OK;
  };
  p.tuning = true;
  float best = 1e30f;
  int bR = 1, bG = 0, bT = 4, bF = 0;
  for (int r : {1, 2, 4}) {
    p.R = r;
    HIPCHK(batch_prepare(h->stream, h));
    for (int g : grids) {
      p.grid_stage = g;
      float ms = 0.f;
      if ((rc = time_it(true, ms))) break;
      if (ms < best) {
        best = ms;
        bR = r;
        bG = g;
      }
    }
    if (rc) break;
  }
  best = 1e30f;
  int bS = 0;
  p.stack = 0;
  for (int ty : {4, 8, 12, 16}) {
    if (rc) break;
    p.TY = ty;
    HIPCHK(batch_prepare(h->stream, h));
    for (int g : grids) {
      p.grid_filter = g;
      float ms = 0.f;
      if ((rc = time_it(false, ms))) break;
      if (ms < best) {
        best = ms;
        bT = ty;
        bF = g;
      }
    }
  }
  // column-march stacks of four waves (solves without holes)
  bool holes = false;
  for (int k = 0; k < h->nblk; k++) holes = holes || h->b[k].has_holes;
  for (int st : {4, 8, 16}) {
    if (rc || holes) break;
    p.stack = st;
    HIPCHK(batch_prepare(h->stream, h));
    for (int g : grids) {
      p.grid_filter = g;
      float ms = 0.f;
      if ((rc = time_it(false, ms))) break;
      if (ms < best) {
        best = ms;
        bS = st;
        bF = g;
      }
    }
  }
  p.stack = 0;
  p.tuning = false;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  batch_release(h);
  if (rc) return rc;
  p.R = bR;
  p.grid_stage = bG;
  p.TY = bT;
  p.stack = bS;
  p.grid_filter = bF;
  p.tuned = true;
  return LEE_OK;
}

// lee_solve for a batch: per chunk of steps, the host schedule goes to the
// device, the step counter to 0, then the captured step is replayed.
int batch_solve(lee_handle *h, int nsteps) {
  BatchPlan &p = h->bp;
  int rc = batch_autotune(h);
  if (rc) return rc;
  if (!p.ready) {
    HIPCHK(batch_prepare(h->stream, h));
  }
  const int chunk_max = 2048;
  rc = batch_reserve(h, (size_t)std::min(nsteps, chunk_max));
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int done = 0; done < nsteps;) {
    const int m = std::min(nsteps - done, chunk_max);
    // next pinned staging buffer: wait until the device has read it
    const int u = h->bp_host_cur;
    h->bp_host_cur ^= 1;
    if (h->bp_host_pending[u]) HIPCHK(hipEventSynchronize(h->bp_host_ev[u]));
    h->bp_host_pending[u] = false;
    const size_t cnt = (size_t)m * h->nblk * h->cfg.rk_stages;
    if (h->bp_host_cap[u] < cnt) {
      if (h->bp_host[u]) HIPCHK(hipHostFree(h->bp_host[u]));
      h->bp_host[u] = nullptr;
      h->bp_host_cap[u] = 0;
      HIPCHK(hipHostMalloc((void **)&h->bp_host[u], sizeof(BatchSched) * cnt, hipHostMallocDefault));
      h->bp_host_cap[u] = cnt;
    }
    if (!h->bp_host_ev[u]) HIPCHK(hipEventCreateWithFlags(&h->bp_host_ev[u], hipEventDisableTiming));
    batch_schedule(h, m, h->bp_host[u]);
    HIPCHK(hipMemcpyAsync(p.sched, h->bp_host[u], sizeof(BatchSched) * cnt, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(p.counter, h->bp_zero, sizeof(h->bp_zero), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipEventRecord(h->bp_host_ev[u], h->stream));
    h->bp_host_pending[u] = true;
    const bool graph = p.use_graph && !h->ktiming;
    if (graph) {
      rc = batch_capture(h);
      if (rc) return rc;
    }
    for (int k = 0; k < m; k++) {
      if (graph) {
        HIPCHK(hipGraphLaunch(p.exec, h->stream));
      } else {
        rc = enqueue_batch_step(h);
        if (rc) return rc;
      }
    }
    done += m;
  }
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  return LEE_OK;
}

// Tile-shape autotuner of the fused kernels.  For each block, every stage
// tile height (rows per thread R = 1, 2, 4) and filter tile height (4, 8,
// 12, 16; one workgroup per tile or the persistent pipelined form with 2, 4
// or 8 workgroups per CU) is timed with HIP events on the block itself, in
// buffers that carry no state between steps (stage: input Y2, accumulator
// Y3, output Y1; filter: Y1 -> Y2); the fastest is kept and cached per (ni,
// nj, nk, device, precision, mean kind, holes, filter order, carried w')
// for the life of the process.  The cache is shared by every handle of the
// process and guarded by a mutex, so handles on different host threads may
// tune concurrently (two threads may both time one key; the first result
// stored wins).
struct TuneKey {
  int ni, nj, nk, device, prec, meanf, holes, filter_order, carry_w;
  bool operator==(const TuneKey &o) const {
    return ni == o.ni && nj == o.nj && nk == o.nk && device == o.device && prec == o.prec && meanf == o.meanf &&
           holes == o.holes && filter_order == o.filter_order && carry_w == o.carry_w;
  }
};
struct TuneVal {
  int r, fty, fgrid, mc, ml;
};
std::vector<std::pair<TuneKey, TuneVal>> g_tune_cache;
std::mutex g_tune_mutex;

TuneKey tune_key(const lee_handle *h, const Block &b) {
  const bool d3 = h->cfg.num_dim == 3;
  return TuneKey{b.ni, b.nj, b.nk, h->device, h->cfg.precision, d3 ? 0 : (b.uniform_mean ? 0 : 1),
                 d3 ? 0 : (b.has_holes ? 1 : 0), h->cfg.filter_order, b.carry_w ? 1 : 0};
}

bool tune_lookup(const TuneKey &key, Block &b) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  for (const auto &e : g_tune_cache)
    if (e.first == key) {
      b.tune_r = e.second.r;
      b.tune_fty = e.second.fty;
      b.tune_fgrid = e.second.fgrid;
      b.tune_mc = e.second.mc;
      b.tune_ml = e.second.ml;
      b.tuned = true;
      return true;
    }
  return false;
}

void tune_store(const TuneKey &key, const Block &b) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  for (const auto &e : g_tune_cache)
    if (e.first == key) return;
  g_tune_cache.push_back({key, {b.tune_r, b.tune_fty, b.tune_fgrid, b.tune_mc, b.tune_ml}});
}

int autotune_block(lee_handle *h, Block &b) {
  const TuneKey key = tune_key(h, b);
  if (tune_lookup(key, b)) return LEE_OK;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int reps = 10;
  const StageScalars sc{1.0, 1.0, 0.0, 0.0, 0.0};
  // best of three timings of `reps` launches (after two untimed ones):
  // candidates a few per cent apart are then told apart reliably
  auto time_ms = [&](bool stage, float &ms) -> int {
    for (int k = 0; k < 2; k++) HIPCHK(stage ? fused_stage(h->stream, h, b, -1, sc) : fused_filter(h->stream, h, b, 1, 2));
    ms = 1e30f;
    for (int trial = 0; trial < 3; trial++) {
      HIPCHK(hipEventRecord(e0, h->stream));
      for (int k = 0; k < reps; k++)
        HIPCHK(stage ? fused_stage(h->stream, h, b, -1, sc) : fused_filter(h->stream, h, b, 1, 2));
      HIPCHK(hipEventRecord(e1, h->stream));
      HIPCHK(hipEventSynchronize(e1));
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, t);
    }
    return LEE_OK;
  };
  float best = 1e30f;
  int best_r = b.tune_r;
  const bool wide_fd = h->cfg.fd_order == 11, wide_flt = h->cfg.filter_order == 11;
  for (int r : {1, 2, 4}) {
    if (wide_fd && r != 1) continue;  // the 11-point stage runs on tiles of 4 rows
    b.tune_r = r;
    float ms = 0.f;
    if (int rc = time_ms(true, ms)) return rc;
    if (ms < best) {
      best = ms;
      best_r = r;
    }
  }
  b.tune_r = best_r;
  best = 1e30f;
  int best_f = b.tune_fty, best_g = 0;
  b.tune_mc = b.tune_ml = 0;  // tile forms first
  if (h->cfg.filter_order > 0) {
    // tile height, then one launch per tile or the persistent pipelined
    // form with 2, 4 or 8 workgroups per CU
    // a height whose tile count just fits the resident workgroups saves a
    // second round: 12 rows are offered besides 4, 8 and 16
    for (int ty : {4, 8, 12, 16}) {
      for (int pw : {0, 2, 4, 8}) {
        if (pw > 0 && b.cus <= 0) continue;
        if (wide_flt && (ty > 8 || pw > 0)) continue;  // SFO11P: tiles of 4 or 8 rows, one workgroup each
        b.tune_fty = ty;
        b.tune_fgrid = pw * b.cus;
        float ms = 0.f;
        if (int rc = time_ms(false, ms)) return rc;
        if (ms < best) {
          best = ms;
          best_f = ty;
          best_g = b.tune_fgrid;
        }
      }
    }
  }
  b.tune_fty = best_f;
  b.tune_fgrid = best_g;
  // column-march forms: single waves of 8 or 16 rows (blocks without
  // holes), stacks of four waves of 4, 8 or 16 rows or of eight of 8 rows
  int best_mc = 0, best_ml = 0;
  if (h->cfg.filter_order > 0 && !wide_flt) {
    static const int forms[][2] = {{1, 8}, {1, 16}, {4, 4}, {4, 8}, {4, 16}, {8, 8}};
    for (const auto &fm : forms) {
      if (fm[0] > 1 || !b.has_holes) {
        const int mc = fm[0], ml = fm[1];
        b.tune_mc = mc;
        b.tune_ml = ml;
        float ms = 0.f;
        if (int rc = time_ms(false, ms)) return rc;
        if (ms < best) {
          best = ms;
          best_mc = mc;
          best_ml = ml;
        }
      }
    }
  }
  b.tune_mc = best_mc;
  b.tune_ml = best_ml;
  b.tuned = true;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  tune_store(key, b);
  return LEE_OK;
}

// 3-D blocks: the stage kernel has one shape; the plane pass of the filter
// is timed in the tile form and as stacks of four waves of 4, 8 or 16 rows
// (Y1 -> Y2 through Y3, buffers that carry no state between steps), cached
// per (ni, nj, nk, precision).
int autotune_block3d(lee_handle *h, Block &b) {
  const TuneKey key = tune_key(h, b);
  if (tune_lookup(key, b)) return LEE_OK;
  int best_mc = 0, best_ml = 0;
  if (h->cfg.filter_order > 0) {
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    static const int forms[][2] = {{1, 8}, {1, 16}, {0, 0}, {4, 4}, {4, 8}, {4, 16}};
    for (const auto &fm : forms) {
      b.tune_mc = fm[0];
      b.tune_ml = fm[1];
      for (int k = 0; k < 2; k++) HIPCHK(fused_filter3d(h->stream, h, b, 1, 2, 3));
      float ms = 1e30f;  // best of three timings of 5 launches
      for (int trial = 0; trial < 3; trial++) {
        HIPCHK(hipEventRecord(e0, h->stream));
        for (int k = 0; k < 5; k++) HIPCHK(fused_filter3d(h->stream, h, b, 1, 2, 3));
        HIPCHK(hipEventRecord(e1, h->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        ms = std::min(ms, t);
      }
      if (ms < best) {
        best = ms;
        best_mc = fm[0];
        best_ml = fm[1];
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  b.tune_mc = best_mc;
  b.tune_ml = best_ml;
  b.tuned = true;
  tune_store(key, b);
  return LEE_OK;
}

int autotune(lee_handle *h) {
  if (is_plain(h)) return LEE_OK;
  const bool second_free = h->nblk == 2 && !h->b[1].tuned;
  for (int ib = 0; ib < h->nblk; ib++) {
    if (h->b[ib].tuned) continue;
    int rc = h->cfg.num_dim == 3 ? autotune_block3d(h, h->b[ib]) : autotune_block(h, h->b[ib]);
    if (rc) return rc;
  }
  // a 2-D pair of blocks with the same kernel instance (an overset
  // background and body grid) takes block 0's forms for both, so that each
  // stage and the filter run as one launch (pair_stage_ok, pair_filter_ok);
  // every form gives the same bits.  Block 1's forms stay as they are when
  // the caller set them, or when block 0's filter form cannot serve it.
  if (second_free && h->cfg.num_dim == 2 && !h->batch) {
    Block &b0 = h->b[0], &b1 = h->b[1];
    if (b0.has_holes == b1.has_holes && b0.uniform_mean == b1.uniform_mean) {
      b1.tune_r = b0.tune_r;
      if (b0.tune_mc > 1) {
        b1.tune_mc = b0.tune_mc;
        b1.tune_ml = b0.tune_ml;
        b1.tune_fgrid = 0;
      }
    }
  }
  return LEE_OK;
}

// solver::solve step prologue (solver.cpp:41-75): CFL at the current dt,
// dt rescaled under FIX_CFL, time_sol += dt.
void advance_clock(lee_handle *h) {
  h->time_step++;
  const double cmax = cfl_max(h);
  double dt = h->dt;
  if (h->cfg.fix_cfl) {
    const double ratio = h->cfg.cfl_or_dt / cmax;
    dt *= ratio;
    h->dt = dt;
  }
  h->time_sol += dt;
}

// ------------------------------------------------------------ adjoint

int adjoint_check(lee_handle *h) {
  if (h->batch) return fail(h, LEE_EINVAL, "the adjoint runs on single handles, not batches");
  if (is_plain(h)) return fail(h, LEE_EINVAL, "the adjoint runs on the fused path");
  // passive scalars (LEE_SCALAR) feed nothing back into s', u', v', p': the
  // adjoint of a functional of p' vanishes on them, and the adjoint runs on
  // the five LEE variables alone (checkpoints keep those five)
  if (h->cfg.fd_order == 11 || h->cfg.filter_order == 11)
    return fail(h, LEE_EINVAL, "the adjoint covers the standard schemes (OA4 derivative, 4th- or 8th-order filter)");
  for (int ib = 0; ib < h->nblk; ib++) {
    // harmonic-wave faces and zones are affine in the state: their linear
    // parts are those of a dirichlet_allzero face and an ambient sponge
    if (h->b[ib].carry_w) return fail(h, LEE_EINVAL, "the adjoint needs w' = 0 in 2-D");
  }
  return LEE_OK;
}

int adjoint_alloc(lee_handle *h, Block &b) {
  const size_t n5 = (size_t)NV * b.n;
  hipStream_t s = h->stream;
  if (is_fp32(h)) {
    if (b.A32[0]) return LEE_OK;
    for (int k = 0; k < 5; k++) {
      HIPCHK(dalloc(&b.A32[k], n5));
      HIPCHK(adjoint_zero(s, h, b, k));
    }
    HIPCHK(dalloc(&b.amu32, (size_t)b.n));
  } else {
    if (b.A64[0]) return LEE_OK;
    for (int k = 0; k < 5; k++) {
      HIPCHK(dalloc(&b.A64[k], n5));
      HIPCHK(adjoint_zero(s, h, b, k));
    }
    HIPCHK(dalloc(&b.amu64, (size_t)b.n));
  }
  HIPCHK(adjoint_zero(s, h, b, -1));
  HIPCHK(dalloc(&b.agrad, (size_t)b.n));
  HIPCHK(adjoint_zero_grad(s, b));
  return LEE_OK;
}

// Copies the step state and the carried aux density of every block to or
// from checkpoint slot `slot`.
int checkpoint_copy(lee_handle *h, int slot, bool save) {
  const size_t ts = is_fp32(h) ? sizeof(float) : sizeof(double);
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    const size_t sb = ((size_t)NV * b.n + b.n) * ts;
    char *base = b.ckpt + (size_t)slot * sb;
    char *y = is_fp32(h) ? (char *)b.Y32[0] : (char *)b.Y64[0];
    char *r = is_fp32(h) ? (char *)b.rho32 : (char *)b.rho64;
    const size_t ys = (size_t)NV * b.n * ts, rs = (size_t)b.n * ts;
    if (save) {
      HIPCHK(hipMemcpyAsync(base, y, ys, hipMemcpyDeviceToDevice, h->stream));
      HIPCHK(hipMemcpyAsync(base + ys, r, rs, hipMemcpyDeviceToDevice, h->stream));
    } else {
      HIPCHK(hipMemcpyAsync(y, base, ys, hipMemcpyDeviceToDevice, h->stream));
      HIPCHK(hipMemcpyAsync(r, base + ys, rs, hipMemcpyDeviceToDevice, h->stream));
    }
  }
  return LEE_OK;
}

// One reversed step on every block: step k of the last objective run
// (dt and time_sol of that step), adjoint state in A[0] and the aux-density
// adjoint in amu on entry and on exit.
// One reversed step.  The boundary update after each forward stage (BCs,
// overset interpolation) couples the blocks, so the reversed stages run
// block by block between the transposed boundary updates: the filter's
// transpose, then the last stage's boundary update (on the accumulator,
// A[1]), the last stage, the boundary update that produced its input, the
// stage before it, and so on.
int adjoint_step(lee_handle *h, int k) {
  const int S = h->cfg.rk_stages;
  const double dt = h->obj.hist_dt[k], t = h->obj.hist_t[k];
  const bool d3 = h->cfg.num_dim == 3;
  bool any_bnd = h->ov.nrecv > 0;
  for (int ib = 0; ib < h->nblk; ib++) any_bnd |= !h->b[ib].faces.empty();
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    HIPCHK(d3 ? adjoint_filter3d(h->stream, h, b, 0, 1, 2, !b.uniform_mean)
              : adjoint_filter(h->stream, h, b, 0, 1, !b.uniform_mean));
  }
  auto stage_all = [&](const AdjStage &proto) -> int {
    for (int ib = 0; ib < h->nblk; ib++) {
      Block &b = h->b[ib];
      AdjStage a = proto;
      a.grad = b.cfg.source_kind == LEE_SOURCE_FIELD;
      double s1 = 0.0, s2 = 0.0;
      source_factors(b.cfg, t + h->c_i[proto.src_stage] * dt, s1, s2);
      a.src = s1 * s2;
      HIPCHK(d3 ? adjoint_stage3d(h->stream, h, b, a) : adjoint_stage(h->stream, h, b, a));
    }
    return LEE_OK;
  };
  if (any_bnd) HIPCHK(adjoint_boundary(h->stream, h, 1));
  int prev = 2;
  AdjStage st{STAGE_LAST, 1, -1, -1, 2, -1, h->b_i[S - 1] * dt, 0.0, 0.0, false, S - 1};
  if (int rc = stage_all(st)) return rc;
  for (int i = S - 2; i >= 1; i--) {
    if (any_bnd) HIPCHK(adjoint_boundary(h->stream, h, prev));
    const int out = prev == 2 ? 3 : 2;
    AdjStage m{STAGE_MID, 1, prev, i == S - 2 ? -1 : 4, out, 4, h->b_i[i] * dt, h->a_next[i] * dt, 0.0, false, i};
    if (int rc = stage_all(m)) return rc;
    prev = out;
  }
  if (any_bnd) HIPCHK(adjoint_boundary(h->stream, h, prev));
  AdjStage f{STAGE_FIRST, 1, prev, S > 2 ? 4 : -1, 0, -1, h->b_i[0] * dt, h->a_next[0] * dt, 0.0, false, 0};
  return stage_all(f);
}

int validate(const lee_config *c, const lee_block *blocks, int nblocks) {
  if (!c) return fail(nullptr, LEE_EINVAL, "null config");
  if (c->num_dim != 2 && c->num_dim != 3) return fail(nullptr, LEE_EINVAL, "DIMENSION must be 2 or 3");
  if (c->num_dim == 3 && c->path != LEE_PATH_FUSED)
    return fail(nullptr, LEE_EINVAL, "3-D runs on the fused path (the reference has no 3-D Jacobian)");
  if (c->fd_order != 4 && c->fd_order != 11)
    return fail(nullptr, LEE_EINVAL, "SPATIAL_SCHEME must be STANDARD_CENTRAL ORDER_ACCURACY 4 or FDO11P (11)");
  if (c->filter_order != 0 && c->filter_order != 4 && c->filter_order != 8 && c->filter_order != 11)
    return fail(nullptr, LEE_EINVAL, "filter order must be 0 (NONE), 4, 8 or 11 (SFO11P)");
  if ((c->fd_order == 11 || c->filter_order == 11) && c->num_dim == 3)
    return fail(nullptr, LEE_EINVAL, "the FDO11P and SFO11P schemes run in 2-D");
  if (c->axisym != 0 && c->axisym != 1) return fail(nullptr, LEE_EINVAL, "AXISYMMETRY must be 0 or 1");
  // LEE_SCALAR (input.cpp:113-119): at least one scalar is the reference's
  // own check, made by the deck reader; the arrays here hold at most four
  if (c->num_scalar < 0 || c->num_scalar > LEE_MAX_SCALAR)
    return fail(nullptr, LEE_EINVAL, "NUMBER_SCALAR must be 0 (LEE) to 4");
  if (c->num_scalar > 0 && c->num_dim != 2)
    return fail(nullptr, LEE_EINVAL, "passive scalars (LEE_SCALAR) run in 2-D");
  if (c->axisym && c->num_dim != 2) return fail(nullptr, LEE_EINVAL, "Axisymmetry requires 2-D.");
  if (c->filter_order == 11 && !(c->filter_strength >= 0.0 && c->filter_strength <= 1.0))
    return fail(nullptr, LEE_EINVAL, "SFO11P STRENGTH must lie in [0, 1]");
  if (c->filter_blend < 0.0 || c->filter_blend > 1.0)
    return fail(nullptr, LEE_EINVAL, "The filter-blending parameter is out-of-bound.");
  if (c->rk_stages != 2 && c->rk_stages != 4) return fail(nullptr, LEE_EINVAL, "Unknown temporal scheme.");
  if (c->precision != LEE_FP64 && c->precision != LEE_FP32) return fail(nullptr, LEE_EINVAL, "bad precision");
  if (c->path != LEE_PATH_FUSED && c->path != LEE_PATH_PLAIN) return fail(nullptr, LEE_EINVAL, "bad path");
  if (c->path == LEE_PATH_PLAIN && c->precision != LEE_FP64)
    return fail(nullptr, LEE_EINVAL, "the plain path runs in FP64 only");
  if (c->source_kind < LEE_SOURCE_NONE || c->source_kind > LEE_SOURCE_FIELD)
    return fail(nullptr, LEE_EINVAL, "unknown source");
  if (c->source_kind == LEE_SOURCE_MONOPOLE && !(c->source_params[5] > 0.0))
    return fail(nullptr, LEE_EINVAL, "monopole half-width must be positive");
  // UserInput::check_consistency_wave (input.cpp:630-655) as the reference
  // runs it: its `else` binds to the inner `if`, so an acoustic wave of a
  // known shape ends in "Unknown wave." and an entropy wave of any shape
  // passes to the amplitude check
  if (c->harmonic_wave < LEE_WAVE_NONE || c->harmonic_wave > LEE_WAVE_ENTROPY)
    return fail(nullptr, LEE_EINVAL, "unknown HARMONIC_WAVE type");
  if (c->harmonic_wave == LEE_WAVE_ACOUSTIC)
    return fail(nullptr, LEE_EINVAL,
                c->harmonic_shape == LEE_WAVEFORM_PLANE || c->harmonic_shape == LEE_WAVEFORM_HOMOGENEOUS
                    ? "Unknown wave."
                    : "Unknown spatial distribution of acoustic waves.");
  if (c->harmonic_wave == LEE_WAVE_ENTROPY) {
    if (c->harmonic_amplitude < 0.0 || c->harmonic_amplitude > 1.0)
      return fail(nullptr, LEE_EINVAL, "The wave amplitude should be between 0 and 1.");
    if (!(c->harmonic_wavelength > 0.0)) return fail(nullptr, LEE_EINVAL, "HARMONIC_WAVE WAVELENGTH must be positive");
    if (c->harmonic_direction < 0 || c->harmonic_direction > 2)
      return fail(nullptr, LEE_EINVAL, "HARMONIC_WAVE DIRECTION must be X, Y or Z");
    if (c->harmonic_shape == LEE_WAVEFORM_GAUSSIAN_ROUNDJET && !(c->harmonic_halfwidth > 0.0))
      return fail(nullptr, LEE_EINVAL, "HARMONIC_WAVE HALFWIDTH must be positive");
  }
  if (nblocks < 1 || nblocks > MAX_BLOCKS) return fail(nullptr, LEE_EINVAL, "block count must be 1..4096");
  // an interior row must fit: 2 half + 1 points for each operator
  const int fd_len = c->fd_order == 11 ? 11 : 5;
  const int flt_len = c->filter_order == 11 ? 11 : c->filter_order + 1;
  const int minlen = std::max(fd_len, flt_len);
  for (int ib = 0; ib < nblocks; ib++) {
    const lee_block &b = blocks[ib];
    if (!b.xyz || !b.mean) return fail(nullptr, LEE_EINVAL, "block without coordinates or base state");
    if (c->num_dim == 2 && b.nk != 1) return fail(nullptr, LEE_EINVAL, "2-D blocks need nk = 1");
    if (b.ni < minlen || b.nj < minlen || (c->num_dim == 3 && b.nk < minlen))
      return fail(nullptr, LEE_EINVAL, "block too small for the stencils");
    if ((long long)b.ni * b.nj * b.nk > (1LL << 30)) return fail(nullptr, LEE_EINVAL, "block too large");
    for (int q = 0; q < b.nboundary; q++) {
      const lee_boundary &r = b.boundary[q];
      const int nd[3] = {b.ni, b.nj, b.nk};
      if (r.dir < 0 || r.dir >= c->num_dim || (r.end != 0 && r.end != 1))
        return fail(nullptr, LEE_EINVAL, "boundary with a bad direction or side");
      for (int d = 0; d < 3; d++)
        if (r.is[d] < 0 || r.ie[d] >= nd[d] || r.is[d] > r.ie[d])
          return fail(nullptr, LEE_EINVAL, "boundary box outside the block");
      switch (r.type) {
        case LEE_BC_DIRICHLET_ALLZERO:
        case LEE_BC_NEUMANN:
        case LEE_BC_WALL_SLIP_KINEMATIC:
        case LEE_BC_CENTERLINE_CART_NORM2X:
        case LEE_BC_CENTERLINE_CART_NORM2Y:
        case LEE_BC_CENTERLINE_CART_NORM2Z:
        case LEE_BC_CENTERLINE_AXISYM:
        case LEE_SPONGE_FREUND_AMBIENT:
          break;
        case LEE_BC_DIRICHLET_HARMONICWAVE:
        case LEE_SPONGE_FREUND_HARMONICWAVE: {
          // bc_dirichlet_harmonicwave's own checks (boundary_condition.cpp:266-322)
          const int sh = c->harmonic_shape;
          if (sh != LEE_WAVEFORM_PLANE && sh != LEE_WAVEFORM_HOMOGENEOUS && sh != LEE_WAVEFORM_GAUSSIAN_ROUNDJET)
            return fail(nullptr, LEE_EINVAL, "Unknown type of wave form.");
          if (c->harmonic_wave != LEE_WAVE_ENTROPY || sh == LEE_WAVEFORM_PLANE)
            return fail(nullptr, LEE_EINVAL, "Unknown type of wave for this wave form.");
          break;
        }
        case LEE_BC_DIRICHLET:
        case LEE_SPONGE_FREUND_DIRICHLET:
          // State::prescribe_on_boundary_solution for LEE and LEE_SCALAR (state.cpp:746-754)
          return fail(nullptr, LEE_EINVAL,
                      c->num_scalar > 0
                          ? "Dirichlet boundary for the linearized Euler model with a scalar is not implemented yet."
                          : "Dirichlet boundary for the linearized Euler model is not implemented yet.");
        default:
          return fail(nullptr, LEE_EINVAL, "unsupported boundary type " + std::to_string(r.type));
      }
      if (c->num_dim == 3 && r.type != LEE_SPONGE_FREUND_AMBIENT)
        return fail(nullptr, LEE_EINVAL, "3-D blocks take ambient sponges only");
    }
  }
  return LEE_OK;
}

}  // namespace

extern "C" {

const char *lee_version(void) { return "camille_amd lee 0.1 (gfx950)"; }

const char *lee_last_error(const lee_handle *h) { return h ? h->err.c_str() : g_create_err.c_str(); }

void lee_destroy(lee_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (int ib = 0; ib < h->nblk; ib++) free_block(h->b[ib]);
  batch_release(h);
  dfree(h->obj.dJ);
  dfree(h->bp.sched);
  dfree(h->bp.counter);
  for (int u = 0; u < 2; u++) {
    if (h->bp_host[u]) (void)hipHostFree(h->bp_host[u]);
    if (h->bp_host_ev[u]) (void)hipEventDestroy(h->bp_host_ev[u]);
  }
  dfree(h->ov.rb);
  dfree(h->ov.rl);
  dfree(h->ov.db);
  dfree(h->ov.dl);
  dfree(h->ov.w);
  dfree(h->ov.coef);
  dfree(h->ov.coef5);
  dfree(h->ov.buf);
  if (h->fdw) (void)hipFree(h->fdw);
  if (h->flw) (void)hipFree(h->flw);
  for (auto e : h->evpool) (void)hipEventDestroy(e);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int lee_create(const lee_config *cfg, const lee_block *blocks, int nblocks, const lee_overset *ov,
               lee_handle **out) {
  if (!out) return fail(nullptr, LEE_EINVAL, "null output handle");
  *out = nullptr;
  int rc = validate(cfg, blocks, nblocks);
  if (rc) return rc;
  return create_impl(cfg, 0, blocks, nblocks, ov, false, out);
}

int lee_batch_create(const lee_config *cfgs, const lee_block *blocks, int nsolve, lee_handle **out) {
  if (!out) return fail(nullptr, LEE_EINVAL, "null output handle");
  *out = nullptr;
  if (!cfgs || !blocks || nsolve < 1) return fail(nullptr, LEE_EINVAL, "empty batch");
  int rc = validate(&cfgs[0], blocks, nsolve);
  if (rc) return rc;
  for (int k = 0; k < nsolve; k++) {
    const lee_config &a = cfgs[0], &c = cfgs[k];
    rc = validate(&c, &blocks[k], 1);
    if (rc) return rc;
    if (c.num_dim != a.num_dim || c.gamma != a.gamma || c.fd_order != a.fd_order ||
        c.filter_order != a.filter_order || c.filter_blend != a.filter_blend || c.rk_stages != a.rk_stages ||
        c.sponge_order != a.sponge_order || c.sponge_constant != a.sponge_constant || c.precision != a.precision ||
        c.path != a.path)
      return fail(nullptr, LEE_EINVAL, "batched solves must share the scheme settings (solve " + std::to_string(k) + ")");
  }
  for (int k = 0; k < nsolve; k++)
    for (int q = 0; q < blocks[k].nboundary; q++)
      if (blocks[k].boundary[q].type == LEE_BC_DIRICHLET_HARMONICWAVE ||
          blocks[k].boundary[q].type == LEE_SPONGE_FREUND_HARMONICWAVE)
        return fail(nullptr, LEE_EINVAL, "batched solves take no harmonic-wave boundaries");
  if (cfgs[0].fd_order == 11 || cfgs[0].filter_order == 11)
    return fail(nullptr, LEE_EINVAL, "batched solves take the standard schemes");
  for (int k = 0; k < nsolve; k++)
    if (cfgs[k].num_scalar > 0) return fail(nullptr, LEE_EINVAL, "batched solves carry no passive scalars");
  if (cfgs[0].path != LEE_PATH_FUSED) return fail(nullptr, LEE_EINVAL, "batched solves run on the fused path");
  if (cfgs[0].filter_order == 0) return fail(nullptr, LEE_EINVAL, "batched solves need the filter");
  return create_impl(cfgs, 1, blocks, nsolve, nullptr, true, out);
}

}  // extern "C"

namespace {

// simulation::initialize from the base state onwards, for a multi-block run
// (one config) or a batch of independent solves (one config per block,
// cfg_stride = 1).
int create_impl(const lee_config *cfgs, int cfg_stride, const lee_block *blocks, int nblocks, const lee_overset *ov,
                bool batch, lee_handle **out) {
  const lee_config *cfg = cfgs;
  lee_handle *h = new (std::nothrow) lee_handle();
  if (!h) return fail(nullptr, LEE_ENOMEM, "out of host memory");
  h->cfg = *cfg;
  h->nblk = nblocks;
  h->batch = batch;
  h->b.resize(nblocks);
  for (int ib = 0; ib < nblocks; ib++) {
    h->b[ib].cfg = cfgs[(size_t)ib * cfg_stride];
    h->b[ib].clk_dt = h->b[ib].cfg.fix_cfl ? 1e-6 : h->b[ib].cfg.cfl_or_dt;  // input.cpp:130
  }
  int rc = [&]() -> int {
    HIPCHK(hipGetDevice(&h->device));
    // the allocation-free entry point's staging area on this device; a
    // failed look-up is kept and reported by that entry point only
    (void)noalloc_prepare_device(h->device);
    int cus = 0;
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    for (auto &blk : h->b) blk.cus = cus;
    // a blocking stream: host-blocking copies order with the queued kernels
    HIPCHK(hipStreamCreate(&h->stream));
    HIPCHK(hipEventCreate(&h->ev0));
    HIPCHK(hipEventCreate(&h->ev1));
    // both recorded once, so lee_synchronize can time a stream that has
    // not run lee_solve yet
    HIPCHK(hipEventRecord(h->ev0, h->stream));
    HIPCHK(hipEventRecord(h->ev1, h->stream));
    make_fd_table(h->fd, cfg->fd_order);
    if (cfg->filter_order > 0) make_filter_table(h->flt, cfg->filter_order, cfg->filter_strength);
    // the fused kernels read the 11-point rows from device memory
    auto upload_wide = [&](const OpTable &t, WideOp **dst) -> int {
      const WideOp w = make_wide_op(t);
      HIPCHK(hipMalloc((void **)dst, sizeof(WideOp)));
      HIPCHK(hipMemcpy(*dst, &w, sizeof(WideOp), hipMemcpyHostToDevice));
      return LEE_OK;
    };
    if (!is_plain(h) && cfg->fd_order == 11)
      if (int e = upload_wide(h->fd, &h->fdw)) return e;
    if (!is_plain(h) && cfg->filter_order == 11)
      if (int e = upload_wide(h->flt, &h->flw)) return e;
    // Butcher tableau (RungeKutta::initialize, temporal:33-85)
    if (cfg->rk_stages == 4) {
      h->b_i[0] = 1.0 / 6.0; h->b_i[1] = 1.0 / 3.0; h->b_i[2] = 1.0 / 3.0; h->b_i[3] = 1.0 / 6.0;
      h->c_i[0] = 0.0; h->c_i[1] = 1.0 / 2.0; h->c_i[2] = 1.0 / 2.0; h->c_i[3] = 1.0;
      h->a_next[0] = 1.0 / 2.0; h->a_next[1] = 1.0 / 2.0; h->a_next[2] = 1.0;
    } else {
      h->b_i[0] = 1.0 / 2.0; h->b_i[1] = 1.0 / 2.0;
      h->c_i[0] = 0.0; h->c_i[1] = 1.0;
      h->a_next[0] = 1.0;
    }
    h->dt = cfg->fix_cfl ? 1e-6 : cfg->cfl_or_dt;  // input.cpp:130
    hipStream_t s = h->stream;
    // iblank from the overset tables: holes, then receivers
    // (override_iblank_4holeCells / _4interpolatedCells, overset.cpp:683-790)
    for (int ib = 0; ib < nblocks; ib++) {
      Block &b = h->b[ib];
      b.ni = blocks[ib].ni;
      b.nj = blocks[ib].nj;
      b.nk = cfg->num_dim == 3 ? blocks[ib].nk : 1;
      b.n = b.ni * b.nj * b.nk;
      b.h_iblank.assign(b.n, (int8_t)NOTBLANKED);
    }
    if (ov && ov->nrecv > 0) {
      if (cfg->num_dim == 3) return fail(h, LEE_EINVAL, "overset interpolation is 2-D only");
      if (nblocks > MAX_OV_BLOCKS) return fail(h, LEE_EINVAL, "an overset run has at most 16 blocks");
      for (int q = 0; q < ov->nhole; q++) {
        const int bb = ov->hole_block[q], l = ov->hole_l[q];
        if (bb < 0 || bb >= nblocks || l < 0 || l >= h->b[bb].n) return fail(h, LEE_EINVAL, "bad hole index");
        h->b[bb].h_iblank[l] = (int8_t)BLANKED;
      }
      for (int r = 0; r < ov->nrecv; r++) {
        const int bb = ov->recv_block[r], l = ov->recv_l[r];
        if (bb < 0 || bb >= nblocks || l < 0 || l >= h->b[bb].n) return fail(h, LEE_EINVAL, "bad receiver index");
        h->b[bb].h_iblank[l] = (int8_t)INTERPOLATED;
      }
    }
    for (int ib = 0; ib < nblocks; ib++) {
      Block &b = h->b[ib];
      const lee_block &in = blocks[ib];
      const int n = b.n;
      b.has_holes = std::any_of(b.h_iblank.begin(), b.h_iblank.end(), [](int8_t v) { return v != 0; });
      b.uniform_mean = true;
      for (int v = 0; v < NV; v++) {
        b.umean[v] = in.mean[(size_t)v * n];
        for (int p = 1; p < n && b.uniform_mean; p++)
          if (in.mean[(size_t)v * n + p] != b.umean[v]) b.uniform_mean = false;
      }
      // the axisymmetric term reads the radius from the coefficient fields
      if (cfg->axisym) b.uniform_mean = false;
      HIPCHK(dalloc(&b.xyz, 3 * (size_t)n));
      const int D = cfg->num_dim;
      HIPCHK(dalloc(&b.met, (size_t)D * D * n));
      HIPCHK(dalloc(&b.jac, (size_t)n));
      const int nv = nvars(h);
      HIPCHK(dalloc(&b.mean, nv * (size_t)n));
      HIPCHK(dalloc(&b.mgrad, (size_t)D * nv * n));
      HIPCHK(dalloc(&b.rhobar, (size_t)n));
      HIPCHK(dalloc(&b.rho64, (size_t)n));
      HIPCHK(dalloc(&b.dflux, (size_t)n));
      HIPCHK(dalloc(&b.dsum, (size_t)n));
      for (int k = 0; k < 4; k++) HIPCHK(dalloc(&b.Y64[k], nv * (size_t)n));
      if (is_fp32(h)) {
        HIPCHK(dalloc(&b.rho32, (size_t)n));
        for (int k = 0; k < 4; k++) HIPCHK(dalloc(&b.Y32[k], nv * (size_t)n));
      }
      if (is_plain(h)) {
        HIPCHK(dalloc(&b.func0, (size_t)n));
        HIPCHK(dalloc(&b.func, (size_t)n));
      }
      for (int k = 0; k < 4; k++) HIPCHK(hipMemsetAsync(b.Y64[k], 0, sizeof(double) * nv * n, s));
      if (is_fp32(h))
        for (int k = 0; k < 4; k++) HIPCHK(hipMemsetAsync(b.Y32[k], 0, sizeof(float) * nv * n, s));
      HIPCHK(hipMemcpyAsync(b.xyz, in.xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(b.mean, in.mean, sizeof(double) * nv * n, hipMemcpyHostToDevice, s));
      // rho_bar = (gamma p_bar)^(1/gamma) e^(-s_bar)  (state.cpp:713-716), host libm
      {
        std::vector<double> rb(n);
        const double g = cfg->gamma, gi = 1.0 / g;
        for (int p = 0; p < n; p++)
          rb[p] = std::pow(g * in.mean[(size_t)IP * n + p], gi) * std::exp(-in.mean[(size_t)IS * n + p]);
        HIPCHK(hipMemcpy(b.rhobar, rb.data(), sizeof(double) * n, hipMemcpyHostToDevice));
      }
      // metrics before any hole (simulation.cpp:207)
      if (D == 3) HIPCHK(metrics3d(s, b));
      else HIPCHK(metrics2d(s, b, h->fd));
      b.h_met.resize((size_t)D * D * n);
      b.h_jac.resize(n);
      HIPCHK(hipMemcpyAsync(b.h_met.data(), b.met, sizeof(double) * D * D * n, hipMemcpyDeviceToHost, s));
      HIPCHK(hipMemcpyAsync(b.h_jac.data(), b.jac, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      // operator codes around holes
      if (b.has_holes) {
        b.h_code_fd.resize(2 * (size_t)n);
        std::string err;
        for (int d = 0; d < 2; d++)
          if (!plan_codes(b, h->fd, d, b.h_code_fd.data() + (size_t)d * n, err)) return fail(h, LEE_EHOLES, err);
        HIPCHK(dalloc(&b.iblank, (size_t)n));
        HIPCHK(dalloc(&b.code_fd, 2 * (size_t)n));
        HIPCHK(hipMemcpy(b.iblank, b.h_iblank.data(), n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(b.code_fd, b.h_code_fd.data(), 2 * (size_t)n, hipMemcpyHostToDevice));
        if (cfg->filter_order > 0) {
          b.h_code_flt.resize(2 * (size_t)n);
          for (int d = 0; d < 2; d++)
            if (!plan_codes(b, h->flt, d, b.h_code_flt.data() + (size_t)d * n, err))
              return fail(h, LEE_EHOLES, err);
          HIPCHK(dalloc(&b.code_flt, 2 * (size_t)n));
          HIPCHK(hipMemcpy(b.code_flt, b.h_code_flt.data(), 2 * (size_t)n, hipMemcpyHostToDevice));
        }
        // integral images of the points off the interior rows (or holes): the
        // fused kernels run hole-free interior tiles without per-point checks
        auto integral = [&](const std::vector<int8_t> &codes, int **dst) -> int {
          const int w = b.ni + 1;
          std::vector<int> S((size_t)w * (b.nj + 1), 0);
          for (int j = 0; j < b.nj; j++)
            for (int i = 0; i < b.ni; i++) {
              const int p = j * b.ni + i;
              const int bad = codes[p] != CODE_INTERIOR || codes[(size_t)n + p] != CODE_INTERIOR ||
                              b.h_iblank[p] == BLANKED;
              S[(size_t)(j + 1) * w + i + 1] = bad + S[(size_t)j * w + i + 1] + S[(size_t)(j + 1) * w + i] -
                                               S[(size_t)j * w + i];
            }
          HIPCHK(dalloc(dst, S.size()));
          HIPCHK(hipMemcpy(*dst, S.data(), sizeof(int) * S.size(), hipMemcpyHostToDevice));
          return LEE_OK;
        };
        if (D == 2) {
          if (int e = integral(b.h_code_fd, &b.clean_fd)) return e;
          if (cfg->filter_order > 0)
            if (int e = integral(b.h_code_flt, &b.clean_flt)) return e;
        }
      }
      // mean gradients after the holes exist (simulation.cpp:240, spatial:2250-2271),
      // the scalars' means included
      for (int v = 0; v < nv; v++)
        for (int d = 0; d < D; d++) {
          if (D == 3) HIPCHK(deriv_xyz3(s, b, b.mean + (size_t)v * n, b.dflux, d));
          else HIPCHK(plain_take_derivative_xyz(s, b, h->fd, b.mean + (size_t)v * n, d));
          HIPCHK(hipMemcpyAsync(b.mgrad + (size_t)(D * v + d) * n, b.dflux, sizeof(double) * n,
                                hipMemcpyDeviceToDevice, s));
        }
      // coefficient fields of a non-uniform mean for the fused stage kernels
      if (!b.uniform_mean) {
        const int nmc = D == 3 ? MC3_COUNT : MC_COUNT;
        std::vector<double> g((size_t)D * NV * n), mc((size_t)nmc * n);
        HIPCHK(hipMemcpyAsync(g.data(), b.mgrad, sizeof(double) * g.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const double gam = cfg->gamma, gi = 1.0 / gam;
        for (int p = 0; D == 3 && p < n; p++) {
          const double pb = in.mean[(size_t)IP * n + p];
          const double rb = std::pow(gam * pb, gi) * std::exp(-in.mean[(size_t)IS * n + p]);
          mc[(size_t)MC3_U * n + p] = in.mean[(size_t)IUX * n + p];
          mc[(size_t)MC3_V * n + p] = in.mean[(size_t)IUY * n + p];
          mc[(size_t)MC3_W * n + p] = in.mean[(size_t)IUZ * n + p];
          mc[(size_t)MC3_GP * n + p] = gam * pb;
          mc[(size_t)MC3_IRB * n + p] = 1.0 / rb;
          for (int q = 0; q < 15; q++) mc[(size_t)(MC3_G + q) * n + p] = g[(size_t)q * n + p];
          mc[(size_t)MC3_RP * n + p] = rb * gi / pb;
          mc[(size_t)MC3_RB * n + p] = rb;
        }
        for (int p = 0; D == 2 && p < n; p++) {
          const double pb = in.mean[(size_t)IP * n + p];
          const double rb = std::pow(gam * pb, gi) * std::exp(-in.mean[(size_t)IS * n + p]);
          auto G = [&](int v, int d) { return g[(size_t)(2 * v + d) * n + p]; };
          const double f[MC_COUNT] = {in.mean[(size_t)IUX * n + p], in.mean[(size_t)IUY * n + p], gam * pb, 1.0 / rb,
                                      G(IS, 0), G(IS, 1), G(IUX, 0), G(IUX, 1), G(IUY, 0), G(IUY, 1), G(IP, 0),
                                      G(IP, 1), rb * gi / pb, rb, in.xyz[(size_t)n + p] + 1e-99};
          for (int c = 0; c < MC_COUNT; c++) mc[(size_t)c * n + p] = f[c];
        }
        if (is_fp32(h)) {
          std::vector<float> m32(mc.begin(), mc.end());
          // r + 1e-99 is 0 in FP32 on the axis, where the numerator vanishes
          // (u_r' = ubar_r = 0): the least normal float keeps 0 / r at 0
          for (int p = 0; D == 2 && p < n; p++)
            m32[(size_t)MC_R * n + p] = std::max(m32[(size_t)MC_R * n + p], FLT_MIN);
          HIPCHK(dalloc(&b.mc32, mc.size()));
          HIPCHK(hipMemcpy(b.mc32, m32.data(), sizeof(float) * m32.size(), hipMemcpyHostToDevice));
        } else {
          HIPCHK(dalloc(&b.mc64, mc.size()));
          HIPCHK(hipMemcpy(b.mc64, mc.data(), sizeof(double) * mc.size(), hipMemcpyHostToDevice));
        }
      }
      // sponges and other boundaries (io.cpp:957-1369)
      const bool roundjet = cfg->harmonic_shape == LEE_WAVEFORM_GAUSSIAN_ROUNDJET;
      // the round-jet Gaussian at the points of a box, box-local (k, j, i) order
      auto gaussian_box = [&](const int *lo, const int *hi, std::vector<double> &out) {
        out.clear();
        for (int k = lo[2]; k <= hi[2]; k++)
          for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
              const size_t p = ((size_t)k * b.nj + j) * b.ni + i;
              out.push_back(harmonic_gaussian(*cfg, in.xyz[p], in.xyz[(size_t)n + p], in.xyz[(size_t)2 * n + p]));
            }
      };
      for (int q = 0; q < in.nboundary; q++) {
        const lee_boundary &r = in.boundary[q];
        if (r.type == LEE_SPONGE_FREUND_AMBIENT || r.type == LEE_SPONGE_FREUND_HARMONICWAVE) {
          SpongeZone z;
          for (int d = 0; d < 3; d++) {
            z.is[d] = r.is[d];
            z.ie[d] = r.ie[d];
          }
          z.harmonic = r.type == LEE_SPONGE_FREUND_HARMONICWAVE;
          if (z.harmonic && roundjet) {
            gaussian_box(z.is, z.ie, z.host_hg);
            HIPCHK(dalloc(&z.hg, z.host_hg.size()));
            HIPCHK(hipMemcpy(z.hg, z.host_hg.data(), sizeof(double) * z.host_hg.size(), hipMemcpyHostToDevice));
          }
          compute_sponge_zone(b, in.xyz, z, r.dir, r.end, cfg->sponge_order, cfg->sponge_constant);
          const size_t cnt = z.host_sigma.size();
          HIPCHK(dalloc(&z.sigma64, cnt));
          HIPCHK(hipMemcpy(z.sigma64, z.host_sigma.data(), sizeof(double) * cnt, hipMemcpyHostToDevice));
          if (is_fp32(h)) {
            std::vector<float> f(cnt);
            for (size_t k = 0; k < cnt; k++) f[k] = (float)z.host_sigma[k];
            HIPCHK(dalloc(&z.sigma32, cnt));
            HIPCHK(hipMemcpy(z.sigma32, f.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
          }
          b.zones.push_back(std::move(z));
        } else {
          BcFace f;
          f.type = r.type;
          f.dir = r.dir;
          f.end = r.end;
          for (int d = 0; d < 3; d++) {
            f.is[d] = r.is[d];
            f.ie[d] = r.ie[d];
          }
          const int face = r.end == 0 ? r.is[r.dir] : r.ie[r.dir];
          f.is[r.dir] = f.ie[r.dir] = face;
          if (r.type == LEE_BC_DIRICHLET_HARMONICWAVE && roundjet) {
            gaussian_box(f.is, f.ie, f.host_hg);
            HIPCHK(dalloc(&f.hg, f.host_hg.size()));
            HIPCHK(hipMemcpy(f.hg, f.host_hg.data(), sizeof(double) * f.host_hg.size(), hipMemcpyHostToDevice));
          }
          b.faces.push_back(std::move(f));
        }
      }
      if ((int)b.zones.size() > MAX_ZONES) return fail(h, LEE_EINVAL, "more than 8 sponge zones in a block");
      plan_source(b.cfg, b, in.xyz);
      if (!b.src.host_g.empty()) {
        HIPCHK(dalloc(&b.src.g, b.src.host_g.size()));
        HIPCHK(hipMemcpy(b.src.g, b.src.host_g.data(), sizeof(double) * b.src.host_g.size(),
                         hipMemcpyHostToDevice));
      }
      HIPCHK(hipStreamSynchronize(s));
      plan_cfl(b, in.mean, cfg->gamma);
    }
    if (ov && ov->nrecv > 0) {
      Overset &o = h->ov;
      o.nrecv = ov->nrecv;
      o.stride = ov->coef_stride;
      o.h_rb.assign(ov->recv_block, ov->recv_block + o.nrecv);
      o.h_rl.assign(ov->recv_l, ov->recv_l + o.nrecv);
      o.h_db.assign(ov->donor_block, ov->donor_block + o.nrecv);
      o.h_dl.assign(ov->donor_l, ov->donor_l + o.nrecv);
      o.h_w.assign(ov->widths, ov->widths + 3 * nblocks);
      o.h_coef.assign(ov->coef, ov->coef + (size_t)o.nrecv * o.stride);
      for (int r = 0; r < o.nrecv; r++) {
        const int db = o.h_db[r];
        if (db < 0 || db >= nblocks) return fail(h, LEE_EINVAL, "bad donor block");
        const int w1 = o.h_w[3 * db], w2 = o.h_w[3 * db + 1], w3 = o.h_w[3 * db + 2];
        const int ns = w1 * w2 * w3;
        if (w3 != 1 || ns > o.stride || ns < 1) return fail(h, LEE_EINVAL, "bad donor stencil width");
        const Block &d = h->b[db];
        const int l = o.h_dl[r], i = l % d.ni, j = l / d.ni;
        if (l < 0 || i + w1 - 1 >= d.ni || j + w2 - 1 >= d.nj)
          return fail(h, LEE_EINVAL, "donor stencil leaves its block");  // overset.cpp:536-537
        // initialize_donor (overset.cpp:516-526): c *= 1 / sum(c)
        double *c = o.h_coef.data() + (size_t)r * o.stride;
        double sum = 0.0;
        for (int q = 0; q < ns; q++) sum += c[q];
        sum = 1.0 / sum;
        for (int q = 0; q < ns; q++) c[q] *= sum;
      }
      HIPCHK(dalloc(&o.rb, o.nrecv));
      HIPCHK(dalloc(&o.rl, o.nrecv));
      HIPCHK(dalloc(&o.db, o.nrecv));
      HIPCHK(dalloc(&o.dl, o.nrecv));
      HIPCHK(dalloc(&o.w, 3 * (size_t)nblocks));
      HIPCHK(dalloc(&o.coef, (size_t)o.nrecv * o.stride));
      HIPCHK(dalloc(&o.buf, (size_t)o.nrecv * nvars(h)));
      HIPCHK(hipMemcpy(o.rb, o.h_rb.data(), sizeof(int) * o.nrecv, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(o.rl, o.h_rl.data(), sizeof(int) * o.nrecv, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(o.db, o.h_db.data(), sizeof(int) * o.nrecv, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(o.dl, o.h_dl.data(), sizeof(int) * o.nrecv, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(o.w, o.h_w.data(), sizeof(int) * 3 * nblocks, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(o.coef, o.h_coef.data(), sizeof(double) * o.h_coef.size(), hipMemcpyHostToDevice));
      // ORDER 4 in 2-D everywhere: a receiver-minor copy of the coefficients
      // for the (receiver, variable)-parallel gather
      bool all5 = true;
      for (int b2 = 0; b2 < nblocks; b2++)
        all5 &= o.h_w[3 * b2] == 5 && o.h_w[3 * b2 + 1] == 5 && o.h_w[3 * b2 + 2] == 1;
      if (all5) {
        std::vector<double> t((size_t)25 * o.nrecv);
        for (int r = 0; r < o.nrecv; r++)
          for (int q = 0; q < 25; q++) t[(size_t)q * o.nrecv + r] = o.h_coef[(size_t)r * o.stride + q];
        HIPCHK(dalloc(&o.coef5, t.size()));
        HIPCHK(hipMemcpy(o.coef5, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
      }
    }
    return LEE_OK;
  }();
  if (rc == LEE_OK && batch) {
    for (int ib = 1; ib < nblocks; ib++)
      if (h->b[ib].uniform_mean != h->b[0].uniform_mean) {
        h->err = "batched solves must all have a uniform mean or all a non-uniform one";
        rc = LEE_EINVAL;
      }
    if (rc == LEE_OK) {
      hipError_t e = hipMalloc((void **)&h->bp.counter, sizeof(h->bp_zero));
      if (e == hipSuccess) e = hipMemcpy(h->bp.counter, h->bp_zero, sizeof(h->bp_zero), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        h->err = std::string("HIP error ") + hipGetErrorString(e);
        rc = LEE_EHIP;
      }
    }
  }
  if (rc) {
    g_create_err = h->err;
    lee_destroy(h);
    return rc;
  }
  *out = h;
  return LEE_OK;
}

}  // namespace

extern "C" {

int lee_set_solution(lee_handle *h, int ib, const double *sol) {
  if (!h || !sol || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  const int n = b.n;
  hipStream_t s = h->stream;
  HIPCHK(hipStreamSynchronize(s));
  b.carry_w = false;
  for (int p = 0; h->cfg.num_dim == 2 && p < n; p++)  // in 3-D w' is an active variable
    if (sol[(size_t)IUZ * n + p] != 0.0) {
      b.carry_w = true;
      break;
    }
  if (h->batch && b.carry_w)
    return fail(h, LEE_EINVAL, "batched 2-D solves carry s', u', v', p' only: w' must be zero");
  const int nv = nvars(h);
  HIPCHK(hipMemcpy(b.Y64[0], sol, sizeof(double) * nv * n, hipMemcpyHostToDevice));
  // compute_dependent_variables after a restart (simulation.cpp:178-180)
  HIPCHK(aux_density(s, b, h->cfg.gamma, b.Y64[0], b.rho64));
  if (is_fp32(h)) {
    HIPCHK(cast_state(s, b, nv, b.Y64[0], b.Y32[0], (size_t)nv * n));
    HIPCHK(cast_state(s, b, 1, b.rho64, b.rho32, (size_t)n));
  }
  HIPCHK(hipStreamSynchronize(s));
  h->have_solution = true;
  return LEE_OK;
}

int lee_get_solution(lee_handle *h, int ib, double *sol) {
  if (!h || !sol || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  hipStream_t s = h->stream;
  const size_t cnt = (size_t)nvars(h) * b.n;
  if (is_fp32(h)) {
    HIPCHK(cast_back(s, b, nvars(h), b.Y32[0], b.Y64[3], cnt));
    HIPCHK(hipMemcpyAsync(sol, b.Y64[3], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  } else {
    HIPCHK(hipMemcpyAsync(sol, b.Y64[0], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_set_time(lee_handle *h, int time_step, double time_sol, double dt) {
  if (!h) return LEE_EINVAL;
  h->time_step = time_step;
  h->time_sol = time_sol;
  h->dt = dt;
  for (auto &b : h->b) {
    b.clk_step = time_step;
    b.clk_time = time_sol;
    b.clk_dt = dt;
  }
  return LEE_OK;
}

int lee_get_time(const lee_handle *h, int *time_step, double *time_sol, double *dt) {
  if (!h) return LEE_EINVAL;
  return lee_get_solve_time(h, 0, time_step, time_sol, dt);
}

int lee_get_solve_time(const lee_handle *h, int isolve, int *time_step, double *time_sol, double *dt) {
  if (!h || isolve < 0 || isolve >= h->nblk) return LEE_EINVAL;
  const bool own = h->batch;
  if (time_step) *time_step = own ? h->b[isolve].clk_step : h->time_step;
  if (time_sol) *time_sol = own ? h->b[isolve].clk_time : h->time_sol;
  if (dt) *dt = own ? h->b[isolve].clk_dt : h->dt;
  return LEE_OK;
}

int lee_compute_cfl(lee_handle *h, double *cmax) {
  if (!h || !cmax) return fail(h, LEE_EINVAL, "bad argument");
  *cmax = h->batch ? cfl_max_block(h->b[0], h->b[0].clk_dt, -1e99) : cfl_max(h);
  return LEE_OK;
}

int lee_set_batch_launch(lee_handle *h, int use_graph, int grid_stage, int grid_filter) {
  if (!h || !h->batch) return fail(h, LEE_EINVAL, "not a batch");
  if (grid_stage < LEE_GRID_RESIDENT || grid_filter < LEE_GRID_RESIDENT)
    return fail(h, LEE_EINVAL, "grid sizes must be >= 0 or LEE_GRID_RESIDENT");
  HIPCHK(hipStreamSynchronize(h->stream));
  h->bp.use_graph = use_graph != 0;
  h->bp.grid_stage = grid_stage;
  h->bp.grid_filter = grid_filter;
  h->bp.tuned = true;
  batch_release(h);
  return LEE_OK;
}

int lee_get_batch_launch(lee_handle *h, int *use_graph, int *grid_stage, int *grid_filter) {
  if (!h || !h->batch || !use_graph || !grid_stage || !grid_filter) return fail(h, LEE_EINVAL, "bad argument");
  *use_graph = h->bp.use_graph;
  *grid_stage = h->bp.grid_stage;
  *grid_filter = h->bp.grid_filter;
  return LEE_OK;
}

int lee_time_advance(lee_handle *h) {
  if (!h) return LEE_EINVAL;
  if (!h->have_solution) return fail(h, LEE_ESTATE, "no solution set");
  if (h->batch) {
    int rc = batch_solve(h, 1);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return LEE_OK;
  }
  int rc = enqueue_step(h);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return LEE_OK;
}

int lee_solve(lee_handle *h, int nsteps) {
  if (!h || nsteps < 0) return fail(h, LEE_EINVAL, "bad argument");
  if (!h->have_solution) return fail(h, LEE_ESTATE, "no solution set");
  if (h->batch) return nsteps > 0 ? batch_solve(h, nsteps) : LEE_OK;
  if (nsteps > 0) {
    int rc = autotune(h);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(h->ev0, h->stream));
  for (int k = 0; k < nsteps; k++) {
    advance_clock(h);
    int rc = enqueue_step(h);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(h->ev1, h->stream));
  return LEE_OK;
}

int lee_synchronize(lee_handle *h) {
  if (!h) return LEE_EINVAL;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipEventElapsedTime(&h->last_ms, h->ev0, h->ev1));
  return LEE_OK;
}

int lee_autotune(lee_handle *h) {
  if (!h) return LEE_EINVAL;
  for (int ib = 0; ib < h->nblk; ib++) h->b[ib].tuned = false;
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache.clear();
  }
  if (h->batch) {
    HIPCHK(hipStreamSynchronize(h->stream));
    h->bp.tuned = false;
    batch_release(h);
    int rc = batch_autotune(h);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return LEE_OK;
  }
  int rc = autotune(h);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return LEE_OK;
}

int lee_set_tile_shape(lee_handle *h, int ib, int stage_rows, int filter_rows) {
  if (!h || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  if (stage_rows != 1 && stage_rows != 2 && stage_rows != 4) return fail(h, LEE_EINVAL, "stage rows must be 1, 2 or 4");
  if (filter_rows != 4 && filter_rows != 8 && filter_rows != 12 && filter_rows != 16)
    return fail(h, LEE_EINVAL, "filter tile height must be 4, 8, 12 or 16");
  if (h->cfg.filter_order == 11 && filter_rows > 8)
    return fail(h, LEE_EINVAL, "the SFO11P filter takes tiles of 4 or 8 rows");
  if (h->batch) {  // one launch covers every solve: the shape is the batch's
    HIPCHK(hipStreamSynchronize(h->stream));
    h->bp.R = stage_rows;
    h->bp.TY = filter_rows;
    h->bp.stack = 0;  // a tile height selects the tile forms
    h->bp.tuned = true;
    batch_release(h);
    return LEE_OK;
  }
  h->b[ib].tune_r = stage_rows;
  h->b[ib].tune_fty = filter_rows;
  h->b[ib].tune_mc = h->b[ib].tune_ml = 0;  // a tile height selects the tile forms
  h->b[ib].tuned = true;
  return LEE_OK;
}

int lee_set_filter_grid(lee_handle *h, int ib, int workgroups) {
  if (!h || ib < 0 || ib >= h->nblk || workgroups < 0) return fail(h, LEE_EINVAL, "bad argument");
  if (h->batch) return fail(h, LEE_EINVAL, "a batch launches its filter through the batch launch settings");
  if (h->cfg.filter_order == 11 && workgroups > 0)
    return fail(h, LEE_EINVAL, "the SFO11P filter runs one workgroup per tile");
  h->b[ib].tune_fgrid = (workgroups + 7) / 8 * 8;
  h->b[ib].tune_mc = h->b[ib].tune_ml = 0;
  h->b[ib].tuned = true;
  return LEE_OK;
}

int lee_set_filter_march(lee_handle *h, int ib, int waves, int rows) {
  if (!h || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  const bool off = waves == 0 && rows == 0;
  if (h->cfg.filter_order == 11 && !off) return fail(h, LEE_EINVAL, "the SFO11P filter runs in the tile form");
  if (h->batch) {  // one launch covers every solve: stacks of four waves only
    if (!off && !(waves == 4 && (rows == 4 || rows == 8 || rows == 16)))
      return fail(h, LEE_EINVAL, "a batch takes stacks of 4 waves of 4, 8 or 16 rows, or (0, 0)");
    for (int k = 0; k < h->nblk && !off; k++)
      if (h->b[k].has_holes) return fail(h, LEE_EINVAL, "the column-march filter serves blocks without holes");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->bp.stack = off ? 0 : rows;
    h->bp.tuned = true;
    batch_release(h);
    return LEE_OK;
  }
  const bool ok = (waves == 1 && (rows == 8 || rows == 16)) ||
                  (waves == 4 && (rows == 4 || rows == 8 || rows == 16)) || (waves == 8 && rows == 8);
  if (!off && !ok)
    return fail(h, LEE_EINVAL,
                "column-march filter: 1 wave of 8 or 16 rows, stacks of 4 waves of 4, 8 or 16 rows or of 8 waves of "
                "8 rows, or (0, 0)");
  if (!off && waves == 1 && h->b[ib].has_holes)
    return fail(h, LEE_EINVAL, "single-wave column-march filter: blocks without holes only (stacks take holes)");
  if (!off && h->cfg.num_dim == 3 && !(waves == 4 || (waves == 1 && (rows == 8 || rows == 16))))
    return fail(h, LEE_EINVAL,
                "a 3-D block takes the single-pass filter (1 wave, tiles of 8 or 16 rows) or the two-pass "
                "filter with its plane pass as stacks of 4 waves of 4, 8 or 16 rows, or (0, 0)");
  h->b[ib].tune_mc = off ? 0 : waves;
  h->b[ib].tune_ml = off ? 0 : rows;
  h->b[ib].tuned = true;
  return LEE_OK;
}

int lee_get_filter_march(lee_handle *h, int ib, int *waves, int *rows) {
  if (!h || ib < 0 || ib >= h->nblk || !waves || !rows) return fail(h, LEE_EINVAL, "bad argument");
  *waves = h->batch ? (h->bp.stack ? 4 : 0) : h->b[ib].tune_mc;
  *rows = h->batch ? h->bp.stack : h->b[ib].tune_ml;
  return LEE_OK;
}

int lee_get_filter_grid(lee_handle *h, int ib, int *workgroups) {
  if (!h || ib < 0 || ib >= h->nblk || !workgroups) return fail(h, LEE_EINVAL, "bad argument");
  *workgroups = h->batch ? 0 : h->b[ib].tune_fgrid;
  return LEE_OK;
}

int lee_get_tile_shape(lee_handle *h, int ib, int *stage_rows, int *filter_rows) {
  if (!h || ib < 0 || ib >= h->nblk || !stage_rows || !filter_rows) return fail(h, LEE_EINVAL, "bad argument");
  *stage_rows = h->batch ? h->bp.R : h->b[ib].tune_r;
  *filter_rows = h->batch ? h->bp.TY : h->b[ib].tune_fty;
  return LEE_OK;
}

int lee_set_kernel_timing(lee_handle *h, int on) {
  if (!h) return LEE_EINVAL;
  HIPCHK(hipStreamSynchronize(h->stream));
  h->ktiming = on != 0;
  h->evused = 0;
  h->evclass.clear();
  return LEE_OK;
}

int lee_get_kernel_timing(lee_handle *h, double *ms, int *count) {
  if (!h || !ms || !count) return fail(h, LEE_EINVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int c = 0; c < 5; c++) {
    ms[c] = 0.0;
    count[c] = 0;
  }
  for (size_t k = 0; k < h->evclass.size(); k++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, h->evpool[2 * k], h->evpool[2 * k + 1]));
    ms[h->evclass[k]] += t;
    count[h->evclass[k]]++;
  }
  h->evused = 0;
  h->evclass.clear();
  return LEE_OK;
}

int lee_last_solve_ms(lee_handle *h, float *ms) {
  if (!h || !ms) return LEE_EINVAL;
  *ms = h->last_ms;
  return LEE_OK;
}

int lee_compute_rhs(lee_handle *h, int ib, const double *y, double time, double *rhs) {
  if (!h || !y || !rhs || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  const int nv = nvars(h);
  const size_t cnt = (size_t)nv * b.n;
  hipStream_t s = h->stream;
  double s1, s2;
  source_factors(b.cfg, time, s1, s2);
  const double he = h->cfg.harmonic_wave == LEE_WAVE_ENTROPY ? harmonic_entropy(h->cfg, time) : 0.0;
  HIPCHK(hipMemcpy(b.Y64[2], y, sizeof(double) * cnt, hipMemcpyHostToDevice));
  if (is_plain(h)) {
    HIPCHK(plain_rhs(s, h, b, b.Y64[2], b.Y64[3], s1, s2, b.dsum, he));
    HIPCHK(hipMemcpyAsync(rhs, b.Y64[3], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  } else {
    StageScalars sc{1.0, 1.0, 0.0, s1, s2, he};
    if (is_fp32(h)) {
      HIPCHK(cast_state(s, b, nv, b.Y64[2], b.Y32[2], cnt));
      HIPCHK(hipMemsetAsync(b.Y32[3], 0, sizeof(float) * cnt, s));
      HIPCHK(h->cfg.num_dim == 3 ? fused_stage3d(s, h, b, -1, sc) : fused_stage(s, h, b, -1, sc));
      HIPCHK(cast_back(s, b, nv, b.Y32[3], b.Y64[3], cnt));
    } else {
      HIPCHK(hipMemsetAsync(b.Y64[3], 0, sizeof(double) * cnt, s));
      HIPCHK(h->cfg.num_dim == 3 ? fused_stage3d(s, h, b, -1, sc) : fused_stage(s, h, b, -1, sc));
    }
    HIPCHK(hipMemcpyAsync(rhs, b.Y64[3], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_apply_filter(lee_handle *h, int ib, double *y) {
  if (!h || !y || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  if (h->cfg.filter_order == 0) return LEE_OK;
  Block &b = h->b[ib];
  const int nv = nvars(h);
  const size_t cnt = (size_t)nv * b.n;
  hipStream_t s = h->stream;
  HIPCHK(hipMemcpy(b.Y64[1], y, sizeof(double) * cnt, hipMemcpyHostToDevice));
  if (is_plain(h)) {
    HIPCHK(plain_filter(s, h, b, b.Y64[1]));
    HIPCHK(hipMemcpyAsync(y, b.Y64[1], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  } else if (is_fp32(h)) {
    HIPCHK(cast_state(s, b, nv, b.Y64[1], b.Y32[1], cnt));
    HIPCHK(cast_state(s, b, nv, b.Y64[1], b.Y32[2], cnt));
    HIPCHK(h->cfg.num_dim == 3 ? fused_filter3d(s, h, b, 1, 2, 3) : fused_filter(s, h, b, 1, 2));
    HIPCHK(cast_back(s, b, nv, b.Y32[2], b.Y64[2], cnt));
    HIPCHK(hipMemcpyAsync(y, b.Y64[2], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  } else {
    HIPCHK(hipMemcpyAsync(b.Y64[2], b.Y64[1], sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
    HIPCHK(h->cfg.num_dim == 3 ? fused_filter3d(s, h, b, 1, 2, 3) : fused_filter(s, h, b, 1, 2));
    HIPCHK(hipMemcpyAsync(y, b.Y64[2], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_take_derivative_xyz(lee_handle *h, int ib, const double *f, int dir_xyz, double *out) {
  if (!h || !f || !out || ib < 0 || ib >= h->nblk || dir_xyz < 0 || dir_xyz >= h->cfg.num_dim)
    return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  hipStream_t s = h->stream;
  HIPCHK(hipMemcpy(b.dsum, f, sizeof(double) * b.n, hipMemcpyHostToDevice));
  if (h->cfg.num_dim == 3) HIPCHK(deriv_xyz3(s, b, b.dsum, b.dflux, dir_xyz));
  else HIPCHK(plain_take_derivative_xyz(s, b, h->fd, b.dsum, dir_xyz));
  HIPCHK(hipMemcpyAsync(out, b.dflux, sizeof(double) * b.n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_update_boundary(lee_handle *h, double *const *ys) {
  if (!h || !ys) return fail(h, LEE_EINVAL, "bad argument");
  hipStream_t s = h->stream;
  for (int ib = 0; ib < h->nblk; ib++)
    HIPCHK(hipMemcpy(h->b[ib].Y64[2], ys[ib], sizeof(double) * nvars(h) * h->b[ib].n, hipMemcpyHostToDevice));
  const double he = h->cfg.harmonic_wave == LEE_WAVE_ENTROPY ? harmonic_entropy(h->cfg, h->time_sol) : 0.0;
  for (int ib = 0; ib < h->nblk; ib++)
    if (!h->b[ib].faces.empty())
      HIPCHK(apply_bcs(s, h, h->b[ib], h->b[ib].Y64[2], nullptr, false, nullptr, nullptr, he));
  HIPCHK(overset_interpolate(s, h, false, 2, false));
  for (int ib = 0; ib < h->nblk; ib++)
    HIPCHK(hipMemcpyAsync(ys[ib], h->b[ib].Y64[2], sizeof(double) * nvars(h) * h->b[ib].n, hipMemcpyDeviceToHost,
                          s));
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_get_metrics(lee_handle *h, int ib, double *metrics, double *jac) {
  if (!h || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  const size_t n = b.n;
  if (metrics) {
    // metrics[l][i] for l, i < 3; a 2-D block fills the upper 2x2
    std::fill(metrics, metrics + 9 * n, 0.0);
    if (h->cfg.num_dim == 3) {
      std::copy(b.h_met.begin(), b.h_met.end(), metrics);
    } else {
      const int map[4] = {0, 1, 3, 4};
      for (int q = 0; q < 4; q++)
        std::copy(b.h_met.begin() + q * n, b.h_met.begin() + (q + 1) * n, metrics + map[q] * n);
    }
  }
  if (jac) std::copy(b.h_jac.begin(), b.h_jac.end(), jac);
  return LEE_OK;
}

int lee_get_iblank(lee_handle *h, int ib, int *iblank) {
  if (!h || !iblank || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  for (int p = 0; p < h->b[ib].n; p++) iblank[p] = h->b[ib].h_iblank[p];
  return LEE_OK;
}

int lee_get_mean_gradient(lee_handle *h, int ib, double *grad) {
  if (!h || !grad || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  HIPCHK(hipMemcpy(grad, b.mgrad, sizeof(double) * h->cfg.num_dim * nvars(h) * b.n, hipMemcpyDeviceToHost));
  return LEE_OK;
}

int lee_get_sponge(lee_handle *h, int ib, double *sigma) {
  if (!h || !sigma || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  std::fill(sigma, sigma + b.n, 0.0);
  for (const auto &z : b.zones)
    for (int k = 0; k < z.nz[2]; k++)
      for (int j = 0; j < z.nz[1]; j++)
        for (int i = 0; i < z.nz[0]; i++)
          sigma[(size_t)(j + z.is[1]) * b.ni + (i + z.is[0])] += z.host_sigma[((size_t)k * z.nz[1] + j) * z.nz[0] + i];
  return LEE_OK;
}

int lee_set_source_field(lee_handle *h, int ib, const double *q) {
  if (!h || !q || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  if (b.cfg.source_kind != LEE_SOURCE_FIELD) return fail(h, LEE_EINVAL, "the block's source is not LEE_SOURCE_FIELD");
  HIPCHK(hipStreamSynchronize(h->stream));
  if (!b.src.g) HIPCHK(dalloc(&b.src.g, (size_t)b.n));
  b.src.host_g.assign(q, q + b.n);
  HIPCHK(hipMemcpy(b.src.g, q, sizeof(double) * b.n, hipMemcpyHostToDevice));
  return LEE_OK;
}

int lee_set_observer(lee_handle *h, int nobs, const int *block, const int *index, const double *weight) {
  if (!h || nobs < 0 || (nobs > 0 && (!block || !index || !weight))) return fail(h, LEE_EINVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    std::vector<int> l;
    std::vector<double> w;
    std::unordered_map<int, int> slot_of;
    b.obs_pos.clear();
    b.obs_slot.clear();
    for (int q = 0; q < nobs; q++) {
      if (block[q] < 0 || block[q] >= h->nblk) return fail(h, LEE_EINVAL, "observer in a bad block");
      if (block[q] != ib) continue;
      if (index[q] < 0 || index[q] >= b.n) return fail(h, LEE_EINVAL, "observer index outside its block");
      auto it = slot_of.find(index[q]);
      if (it == slot_of.end()) {
        it = slot_of.emplace(index[q], (int)l.size()).first;
        l.push_back(index[q]);
        w.push_back(0.0);
      }
      w[it->second] += weight[q];
      b.obs_pos.push_back(q);
      b.obs_slot.push_back(it->second);
    }
    dfree(b.obs_l);
    dfree(b.obs_w);
    dfree(b.obs_buf);
    b.nobs = (int)l.size();
    h->nobs_total = nobs;
    if (b.nobs == 0) continue;
    HIPCHK(dalloc(&b.obs_l, l.size()));
    HIPCHK(dalloc(&b.obs_w, w.size()));
    HIPCHK(dalloc(&b.obs_buf, l.size()));
    HIPCHK(hipMemcpy(b.obs_l, l.data(), sizeof(int) * l.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.obs_w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
  }
  return LEE_OK;
}

int lee_sample_observers(lee_handle *h, double *out) {
  if (!h || !out) return fail(h, LEE_EINVAL, "bad argument");
  std::vector<double> tmp;
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    if (b.nobs == 0) continue;
    HIPCHK(observer_gather(h->stream, h, b, b.obs_buf));
    tmp.resize(b.nobs);
    HIPCHK(hipMemcpyAsync(tmp.data(), b.obs_buf, sizeof(double) * b.nobs, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t e = 0; e < b.obs_pos.size(); e++) out[b.obs_pos[e]] = tmp[b.obs_slot[e]];
  }
  return LEE_OK;
}

int lee_objective_forward(lee_handle *h, int nsteps, int window_start, int ckpt_every, double *J) {
  if (!h || !J || nsteps < 1 || ckpt_every < 1 || window_start < 0 || window_start >= nsteps)
    return fail(h, LEE_EINVAL, "bad argument");
  if (!h->have_solution) return fail(h, LEE_ESTATE, "no solution set");
  int rc = adjoint_check(h);
  if (rc) return rc;
  rc = autotune(h);
  if (rc) return rc;
  Objective &o = h->obj;
  o.ready = false;
  o.nsteps = nsteps;
  o.window = window_start;
  o.every = ckpt_every;
  o.ncheck = (nsteps + ckpt_every - 1) / ckpt_every;
  const size_t ts = is_fp32(h) ? sizeof(float) : sizeof(double);
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    const size_t need = (size_t)(o.ncheck + 1) * ((size_t)NV * b.n + b.n) * ts;
    if (b.ckpt_bytes < need) {
      HIPCHK(hipStreamSynchronize(h->stream));
      dfree(b.ckpt);
      b.ckpt_bytes = 0;
      HIPCHK(dalloc(&b.ckpt, need));
      b.ckpt_bytes = need;
    }
  }
  if (!o.dJ) HIPCHK(dalloc(&o.dJ, 1));
  HIPCHK(hipMemsetAsync(o.dJ, 0, sizeof(double), h->stream));
  o.ck_step.assign(o.ncheck, 0);
  o.ck_time.assign(o.ncheck, 0.0);
  o.ck_dt.assign(o.ncheck, 0.0);
  o.hist_dt.assign(nsteps, 0.0);
  o.hist_t.assign(nsteps, 0.0);
  for (int k = 0; k < nsteps; k++) {
    if (k % ckpt_every == 0) {
      const int c = k / ckpt_every;
      rc = checkpoint_copy(h, c, true);
      if (rc) return rc;
      o.ck_step[c] = h->time_step;
      o.ck_time[c] = h->time_sol;
      o.ck_dt[c] = h->dt;
    }
    advance_clock(h);
    o.hist_dt[k] = h->dt;
    o.hist_t[k] = h->time_sol;
    rc = enqueue_step(h);
    if (rc) return rc;
    if (k + 1 > window_start)
      for (int ib = 0; ib < h->nblk; ib++) HIPCHK(observer_energy(h->stream, h, h->b[ib], o.dJ));
  }
  rc = checkpoint_copy(h, o.ncheck, true);
  if (rc) return rc;
  o.end_step = h->time_step;
  o.end_time = h->time_sol;
  o.end_dt = h->dt;
  HIPCHK(hipMemcpyAsync(J, o.dJ, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  o.ready = true;
  return LEE_OK;
}

int lee_objective_backward(lee_handle *h, double scale) {
  if (!h) return LEE_EINVAL;
  Objective &o = h->obj;
  if (!o.ready) return fail(h, LEE_ESTATE, "no objective forward run to reverse");
  int rc = adjoint_check(h);
  if (rc) return rc;
  hipStream_t s = h->stream;
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    rc = adjoint_alloc(h, b);
    if (rc) return rc;
    HIPCHK(adjoint_zero(s, h, b, 0));
    HIPCHK(adjoint_zero(s, h, b, -1));
    HIPCHK(adjoint_zero_grad(s, b));
    const size_t need = (size_t)o.every * b.nobs;
    if (b.samp_cap < need) {
      HIPCHK(hipStreamSynchronize(s));
      dfree(b.samp);
      b.samp_cap = 0;
      HIPCHK(dalloc(&b.samp, need));
      b.samp_cap = need;
    }
  }
  for (int c = o.ncheck - 1; c >= 0; c--) {
    const int k0 = c * o.every, k1 = std::min(o.nsteps, k0 + o.every);
    // recompute the segment forwards from its checkpoint, keeping the samples
    rc = checkpoint_copy(h, c, false);
    if (rc) return rc;
    h->time_step = o.ck_step[c];
    h->time_sol = o.ck_time[c];
    h->dt = o.ck_dt[c];
    for (int k = k0; k < k1; k++) {
      advance_clock(h);
      if (h->dt != o.hist_dt[k] || h->time_sol != o.hist_t[k])
        return fail(h, LEE_ESTATE, "clock of the recomputed segment differs from the forward run");
      rc = enqueue_step(h);
      if (rc) return rc;
      if (k + 1 > o.window)
        for (int ib = 0; ib < h->nblk; ib++) {
          Block &b = h->b[ib];
          HIPCHK(observer_gather(s, h, b, b.samp + (size_t)(k - k0) * b.nobs));
        }
    }
    // then backwards through it
    for (int k = k1 - 1; k >= k0; k--) {
      if (k + 1 > o.window)
        for (int ib = 0; ib < h->nblk; ib++) {
          Block &b = h->b[ib];
          HIPCHK(observer_seed(s, h, b, 0, b.samp + (size_t)(k - k0) * b.nobs, scale));
        }
      rc = adjoint_step(h, k);
      if (rc) return rc;
    }
  }
  // the first step's stale aux density is R(y_0)
  for (int ib = 0; ib < h->nblk; ib++)
    if (!h->b[ib].uniform_mean)
      HIPCHK(h->cfg.num_dim == 3 ? adjoint_add_rt_mu3(s, h, h->b[ib], 0) : adjoint_add_rt_mu(s, h, h->b[ib], 0));
  rc = checkpoint_copy(h, o.ncheck, false);
  if (rc) return rc;
  h->time_step = o.end_step;
  h->time_sol = o.end_time;
  h->dt = o.end_dt;
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_get_source_gradient(lee_handle *h, int ib, double *grad_q) {
  if (!h || !grad_q || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  if (!b.agrad) return fail(h, LEE_ESTATE, "no adjoint run");
  HIPCHK(hipMemcpy(grad_q, b.agrad, sizeof(double) * b.n, hipMemcpyDeviceToHost));
  return LEE_OK;
}

int lee_get_state_gradient(lee_handle *h, int ib, double *grad_y0) {
  if (!h || !grad_y0 || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  Block &b = h->b[ib];
  const size_t cnt = (size_t)NV * b.n;
  if (is_fp32(h)) {
    if (!b.A32[0]) return fail(h, LEE_ESTATE, "no adjoint run");
    HIPCHK(cast_back(h->stream, b, NV, b.A32[0], b.A64[0] ? b.A64[0] : b.Y64[3], cnt));
    HIPCHK(hipMemcpyAsync(grad_y0, b.A64[0] ? b.A64[0] : b.Y64[3], sizeof(double) * cnt, hipMemcpyDeviceToHost,
                          h->stream));
  } else {
    if (!b.A64[0]) return fail(h, LEE_ESTATE, "no adjoint run");
    HIPCHK(hipMemcpyAsync(grad_y0, b.A64[0], sizeof(double) * cnt, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return LEE_OK;
}

int lee_update_boundary_adjoint(lee_handle *h, double *const *ys) {
  if (!h || !ys) return fail(h, LEE_EINVAL, "bad argument");
  int rc = adjoint_check(h);
  if (rc) return rc;
  hipStream_t s = h->stream;
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    rc = adjoint_alloc(h, b);
    if (rc) return rc;
    const size_t cnt = (size_t)NV * b.n;
    HIPCHK(hipMemcpy(b.Y64[3], ys[ib], sizeof(double) * cnt, hipMemcpyHostToDevice));
    if (is_fp32(h)) HIPCHK(cast_state(s, b, NV, b.Y64[3], b.A32[0], cnt));
    else HIPCHK(hipMemcpyAsync(b.A64[0], b.Y64[3], sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
  }
  HIPCHK(adjoint_boundary(s, h, 0));
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    const size_t cnt = (size_t)NV * b.n;
    if (is_fp32(h)) HIPCHK(cast_back(s, b, NV, b.A32[0], b.Y64[3], cnt));
    else HIPCHK(hipMemcpyAsync(b.Y64[3], b.A64[0], sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(ys[ib], b.Y64[3], sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return LEE_OK;
}

int lee_compute_rhs_adjoint(lee_handle *h, int ib, const double *kappa, double *out) {
  if (!h || !kappa || !out || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  int rc = adjoint_check(h);
  if (rc) return rc;
  Block &b = h->b[ib];
  rc = adjoint_alloc(h, b);
  if (rc) return rc;
  const size_t cnt = (size_t)NV * b.n;
  hipStream_t s = h->stream;
  double *stage = b.Y64[3];  // FP64 staging of host data
  HIPCHK(hipMemcpy(stage, kappa, sizeof(double) * cnt, hipMemcpyHostToDevice));
  if (is_fp32(h)) HIPCHK(cast_state(s, b, NV, stage, b.A32[1], cnt));
  else HIPCHK(hipMemcpyAsync(b.A64[1], stage, sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
  HIPCHK(adjoint_zero(s, h, b, 2));
  AdjStage st{STAGE_MID, 1, -1, -1, 2, -1, 1.0, 0.0, 0.0, false};
  HIPCHK(h->cfg.num_dim == 3 ? adjoint_stage3d(s, h, b, st) : adjoint_stage(s, h, b, st));
  if (is_fp32(h)) HIPCHK(cast_back(s, b, NV, b.A32[2], stage, cnt));
  else HIPCHK(hipMemcpyAsync(stage, b.A64[2], sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(out, stage, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

int lee_apply_filter_adjoint(lee_handle *h, int ib, double *y) {
  if (!h || !y || ib < 0 || ib >= h->nblk) return fail(h, LEE_EINVAL, "bad argument");
  int rc = adjoint_check(h);
  if (rc) return rc;
  Block &b = h->b[ib];
  rc = adjoint_alloc(h, b);
  if (rc) return rc;
  const size_t cnt = (size_t)NV * b.n;
  hipStream_t s = h->stream;
  double *stage = b.Y64[3];
  HIPCHK(hipMemcpy(stage, y, sizeof(double) * cnt, hipMemcpyHostToDevice));
  if (is_fp32(h)) HIPCHK(cast_state(s, b, NV, stage, b.A32[0], cnt));
  else HIPCHK(hipMemcpyAsync(b.A64[0], stage, sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
  HIPCHK(adjoint_zero(s, h, b, 1));
  HIPCHK(h->cfg.num_dim == 3 ? adjoint_filter3d(s, h, b, 0, 1, 2, false) : adjoint_filter(s, h, b, 0, 1, false));
  if (is_fp32(h)) HIPCHK(cast_back(s, b, NV, b.A32[1], stage, cnt));
  else HIPCHK(hipMemcpyAsync(stage, b.A64[1], sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
  // w' is not an adjoint variable in 2-D; it passes through the filter alone
  HIPCHK(hipMemcpyAsync(y, stage, sizeof(double) * cnt, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return LEE_OK;
}

}  // extern "C"
