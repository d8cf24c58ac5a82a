// Internal declarations of camille_amd's LEE time advance (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lee.h"

namespace lee {

constexpr int NV = 5;          // s', u', v', w', p' (src/core/param.h:64-71)
constexpr int MAXV = NV + LEE_MAX_SCALAR;  // and the passive scalars of LEE_SCALAR (state.cpp:336-372)
constexpr int IS = 0, IUX = 1, IUY = 2, IUZ = 3, IP = 4;
constexpr int NOTBLANKED = 0, BLANKED = 1, INTERPOLATED = -1;  // param.h:126-128
constexpr int MAX_ZONES = 8;   // sponge zones per block handled by the fused kernel
constexpr int MAX_BLOCKS = 4096;     // blocks of a handle (solves of a batch)
constexpr int MAX_OV_BLOCKS = 16;    // blocks of an overset run (per-block tables travel as kernel arguments)

// Operator-code convention per point and direction (int8):
//   0            interior row
//   1            zero (hole after hole)
//   2 + k        left-closure row k, anchored k points to the left
//   2 + nb + k   right-closure row k, anchored k points to the right (mirrored)
constexpr int CODE_INTERIOR = 0, CODE_ZERO = 1, CODE_LEFT = 2;
constexpr int STAGE_FIRST = 0, STAGE_MID = 1, STAGE_LAST = 2;  // RK stage kinds
// Mean-flow coefficient fields of a non-uniform base state, in the bulk
// type: u, v, gamma p, 1/rho, d(s, u, v, p)/d(x, y), rho gi / p, rho, and
// the radius r + 1e-99 of an axisymmetric run.
enum { MC_U, MC_V, MC_GP, MC_IRB, MC_GSX, MC_GSY, MC_GUX, MC_GUY, MC_GVX, MC_GVY, MC_GPX, MC_GPY, MC_RP, MC_RB,
       MC_R, MC_COUNT };
// 3-D: u, v, w, gamma p, 1/rho, d(s, u, v, w, p)/d(x, y, z) (5 x 3), rho gi / p, rho.
enum { MC3_U, MC3_V, MC3_W, MC3_GP, MC3_IRB, MC3_G, MC3_RP = MC3_G + 15, MC3_RB, MC3_COUNT };

// One 1-D operator (interior stencil + closure rows), reference tables of
// spatial_discretization.cpp:28-224.
struct OpTable {
  int half;            // half width of the interior stencil
  int nb;              // closure rows
  int sb;              // entries per closure row
  double sign_right;   // -1 for derivatives, +1 for filters
  double st[11];       // interior coefficients, index s + half
  double bnd[5][9];    // closure rows
};

struct SpongeZone {    // one Freund buffer zone (ambient or harmonic-wave target)
  int is[3], ie[3];
  int nz[3];
  double *sigma64 = nullptr;  // device, zone-local (k, j, i) order
  float *sigma32 = nullptr;
  std::vector<double> host_sigma;
  // sponge_freund_harmonicwave: the zone relaxes s' to e(t) g(x) (and the
  // other variables to 0); g is 1 for a homogeneous wave (hg null) or the
  // round-jet Gaussian per zone point
  bool harmonic = false;
  double *hg = nullptr;
  std::vector<double> host_hg;
};

struct BcFace {        // one boundary-condition face (BCs other than sponges)
  int type, dir, end;
  int is[3], ie[3];    // face box (dir index collapsed to the face)
  double *hg = nullptr;  // dirichlet_harmonicwave: g per face point (null: 1)
  std::vector<double> host_hg;
};

struct SourceBox {     // spatial factor of a physical source over its support
  int i0 = 0, i1 = -1, j0 = 0, j1 = -1;  // inclusive box; empty when i1 < i0
  int k0 = 0, k1 = 0;
  double *g = nullptr;                    // device (k-j-i order over the box)
  std::vector<double> host_g;
};

// Per-block device data.
struct Block {
  int ni = 0, nj = 0, nk = 1, n = 0;
  bool has_holes = false;      // any non-NOTBLANKED point
  bool uniform_mean = false;   // every mean field spatially constant
  double umean[NV] = {0};      // the constants when uniform
  // geometry (FP64)
  double *xyz = nullptr;       // 3n, kept for init
  double *met = nullptr;       // 2-D 4n: xi_x, xi_y, eta_x, eta_y; 3-D 9n: m[l][i] at (3 l + i) n
  double *jac = nullptr;       // n
  int8_t *iblank = nullptr;    // n
  int8_t *code_fd = nullptr;   // 2n: xi codes then eta codes (holes only)
  int8_t *code_flt = nullptr;  // 2n
  int *clean_fd = nullptr;     // (ni+1)(nj+1) integral images of the points whose FD /
  int *clean_flt = nullptr;    // filter row is not the interior one or that are holes
  std::vector<int8_t> h_iblank, h_code_fd, h_code_flt;
  // base state (FP64): mean 5n, mean gradient 5*2n, rho_bar n
  double *mean = nullptr, *mgrad = nullptr, *rhobar = nullptr;
  // non-uniform mean: MC_COUNT coefficient fields in the bulk type
  double *mc64 = nullptr;
  float *mc32 = nullptr;
  // aux density rho' carried between steps (stale after the filter)
  double *rho64 = nullptr;
  float *rho32 = nullptr;
  // state buffers: Y[0] step state, Y[1] accumulator, Y[2], Y[3] stage inputs
  double *Y64[4] = {nullptr, nullptr, nullptr, nullptr};
  float *Y32[4] = {nullptr, nullptr, nullptr, nullptr};
  bool carry_w = false;        // 2-D with a non-zero w' (sponge + filter only)
  // tile shape of the fused kernels: stage rows per thread (1, 2, 4; tile
  // height 4 R) and filter tile height (4, 8, 16; FP64 4); set by the autotuner
  int tune_r = 1, tune_fty = 4;
  int tune_fgrid = 0;          // filter: 0 one workgroup per tile, else persistent pipelined grid
  int tune_mc = 0, tune_ml = 0;  // filter column-march form (no holes): waves per stack, rows per wave; 0 off
  int cus = 0;                 // compute units of the device
  bool tuned = false;
  // plain-path scratch (FP64): dflux, sum, rhs (5n), func0, func, interim (5n), ynp1 (5n)
  double *dflux = nullptr, *dsum = nullptr, *rhs = nullptr, *func0 = nullptr, *func = nullptr;
  double *interim = nullptr, *ynp1 = nullptr, *sol = nullptr;
  // boundaries
  std::vector<SpongeZone> zones;
  std::vector<BcFace> faces;
  SourceBox src;
  // CFL: distinct (X, J) pairs that can attain the maximum (see lee_plan.cpp)
  std::vector<std::pair<double, double>> cfl_cand;
  // host copies used by init checks
  std::vector<double> h_met, h_jac;
  // discrete adjoint (allocated on first use): A[0] adjoint of the step
  // state, A[1] of the unfiltered state / accumulator, A[2], A[3] of the
  // stage inputs, A[4] running adjoint of y_n; adjoint of the stale aux
  // density; source-field gradient (FP64)
  double *A64[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  float *A32[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double *amu64 = nullptr;
  float *amu32 = nullptr;
  double *agrad = nullptr;
  int adj_r = 2;               // rows per thread of the adjoint stage kernel (1 or 2)
  // observer points of the functional (device) and their weights
  int nobs = 0;
  int *obs_l = nullptr;
  double *obs_w = nullptr;
  // Observer points of this block, duplicates merged: obs_l/obs_w hold each
  // distinct point once with the sum of its weights (J and the adjoint seed
  // then update every point from one thread).  The caller's entry obs_pos[e]
  // samples the distinct point obs_slot[e].
  std::vector<int> obs_pos;
  std::vector<int> obs_slot;
  double *obs_buf = nullptr;   // device samples of the current state
  // checkpoints of an objective run: slots of (state 5n + aux density n)
  // in the bulk type; the last slot holds the state at the end of the run
  char *ckpt = nullptr;
  size_t ckpt_bytes = 0;
  double *samp = nullptr;      // observer samples of one checkpoint segment
  size_t samp_cap = 0;
  // batched ensembles: every block is an independent solve with its own
  // settings (source, CFL or dt) and clock
  lee_config cfg{};
  int clk_step = 0;
  double clk_time = 0.0, clk_dt = 1e-6;
};

// The optimized 11-point operators (FDO11P, SFO11P) as the fused kernels
// read them from device memory: the OpTable rows with each closure row's
// own length.
struct WideOp {
  double st[11];       // interior row, index s + 5
  double bnd[5][9];    // closure row k, anchored k points to the left
  int len[5];          // entries of closure row k
  int nb;              // closure rows
  double sign_right;   // -1 derivative, +1 filter
};
WideOp make_wide_op(const OpTable &t);

// One reversed RK stage: buffer indices into Block::A (-1: none).
struct AdjStage {
  int kind;                    // 0 first, 1 middle, 2 last stage
  int lacc, lin, lyin, lout, lyout;
  double hb, ha, src;          // b_i h, a_i h, source time factor s1 s2
  bool grad;                   // accumulate the source-field gradient
  int src_stage = 0;           // RK stage index whose time gives src
};

// Bookkeeping of the last lee_objective_forward.
struct Objective {
  bool ready = false;
  int nsteps = 0, window = 0, every = 1, ncheck = 0;
  std::vector<int> ck_step;            // clock at each checkpoint
  std::vector<double> ck_time, ck_dt;
  std::vector<double> hist_dt, hist_t; // per step: dt and time_sol after the increment
  int end_step = 0;
  double end_time = 0.0, end_dt = 0.0;
  double *dJ = nullptr;                // device accumulator
};

// Per-step scalars of one solve in a batched ensemble.
struct BatchSched {
  double dt, s1, s2;
};

// Device tables of a batched ensemble: one persistent launch per RK stage
// and one per filter over the tiles of all solves, replayed from a HIP graph.
struct BatchPlan {
  bool ready = false;          // tables match the tile shape and precision
  bool tuned = false;
  bool tuning = false;         // tables address scratch buffers only (autotuner)
  int R = 1, TY = 4;           // stage rows per thread, filter tile height
  int stack = 0;               // filter: rows per wave of four-wave column-march stacks (0: tiles)
  int grid_stage = 0, grid_filter = 0;  // persistent grid sizes (0: one workgroup per tile)
  int res_stage = 0, res_filter = 0;    // resident workgroups of the prepared stage and filter forms
  void *stage_args = nullptr;  // device StageArgs<T>[stages][nsolve]
  void *filter_args = nullptr; // device FilterArgs<T>[nsolve]
  int *tile_map_stage = nullptr, *tile_map_filter = nullptr;  // (solve, tile) per list entry
  int total_stage = 0, total_filter = 0;
  BatchSched *sched = nullptr;  // device [step][solve][stage]
  size_t sched_steps = 0;
  int *counter = nullptr;      // device: [0] step index within the schedule, [1] scratch,
                               // [32, 288) work queues of the persistent stage launches
  bool use_graph = true;
  hipGraphExec_t exec = nullptr;
};

struct Overset {
  int nrecv = 0;
  int *rb = nullptr, *rl = nullptr, *db = nullptr, *dl = nullptr;  // device
  int *w = nullptr;            // device widths 3 per block
  double *coef = nullptr;      // device normalised coefficients nrecv * stride
  double *coef5 = nullptr;     // when every stencil is 5 x 5: coefficient q of receiver r at q * nrecv + r
  int stride = 0;
  double *buf = nullptr;       // device nrecv * NV
  std::vector<int> h_rb, h_rl, h_db, h_dl, h_w;
  std::vector<double> h_coef;
  double **yptr = nullptr;     // device table of per-block variable pointers (NV per block)
};

}  // namespace lee

struct lee_handle {
  lee_config cfg{};
  int nblk = 0;
  std::vector<lee::Block> b;
  lee::Objective obj;
  int nobs_total = 0;
  bool batch = false;          // independent solves (lee_batch_create)
  lee::BatchPlan bp;
  // pinned host staging of the batch schedule, two buffers used in turn;
  // an event marks when the device has copied each one
  lee::BatchSched *bp_host[2] = {nullptr, nullptr};
  size_t bp_host_cap[2] = {0, 0};
  hipEvent_t bp_host_ev[2] = {nullptr, nullptr};
  bool bp_host_pending[2] = {false, false};
  int bp_host_cur = 0;
  int bp_zero[288] = {};       // the batch counters at the start of a solve
  lee::Overset ov;
  lee::OpTable fd{}, flt{};
  lee::WideOp *fdw = nullptr, *flw = nullptr;  // device copies of FDO11P / SFO11P rows (fused path)
  int nb_fd = 2, nb_flt = 4;
  double b_i[4] = {0}, c_i[4] = {0}, a_next[4] = {0};  // a_next[i] = a_{i+1,i}
  int time_step = 0;
  double time_sol = 0.0, dt = 1e-6;
  bool have_solution = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  std::string err;
  int device = 0;
  // optional per-kernel timing: HIP events around each fused launch,
  // classes 0 first stage, 1 middle stages, 2 last stage, 3 filter, 4 boundary
  bool ktiming = false;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  std::vector<int> evclass;  // one entry per (start, stop) pair
};

namespace lee {

// Variables of a handle's state arrays: NV and the passive scalars.
inline int nvars(const lee_handle *h) { return NV + h->cfg.num_scalar; }

// ---- host-side planning (lee_plan.cpp) ----
void make_fd_table(OpTable &t, int order);
void make_filter_table(OpTable &t, int order, double strength);
// Replays the reference's operator coverage + iblank_correction scan for one
// direction; writes per-point codes.  Returns false (and a message) when a
// closure would cover a hole (the reference asserts, spatial:1794).
bool plan_codes(const Block &b, const OpTable &op, int dir, int8_t *codes, std::string &err);
void compute_sponge_zone(const Block &b, const double *xyz_host, SpongeZone &z, int dir, int end, int order,
                         double constant);
void plan_cfl(Block &b, const double *mean_host, double gamma);
double cfl_max(const lee_handle *h);
double cfl_max_block(const Block &b, double dt, double m);
void plan_source(const lee_config &c, Block &b, const double *xyz_host);
// time factors (s1, s2) of the source at time t: term = (s1 * g) * s2
void source_factors(const lee_config &c, double t, double &s1, double &s2);

// ---- kernels (lee_kernels.hip) ----
struct StageScalars {
  double h;        // dt
  double b;        // b_i
  double a;        // a_{i+1,i}
  double s1, s2;   // source time factors
  double he = 0.0; // harmonic-wave entropy factor e(t) at the stage time
};
hipError_t plain_apply_op(hipStream_t s, const Block &b, const OpTable &op, const int8_t *codes,
                          const double *f, double *df, int dir);
hipError_t plain_take_derivative_xyz(hipStream_t s, const Block &b, const OpTable &op, const double *f,
                                     int dir_xyz);
hipError_t plain_rhs(hipStream_t s, lee_handle *h, const Block &b, const double *y, double *rhs, double s1,
                     double s2, double *rho, double he = 0.0);
hipError_t plain_rk_update(hipStream_t s, const Block &b, double *rhs, double *ynext, double *interim,
                           const double *ycur, double hdt, double bi, double a, bool last, int nv);
hipError_t plain_filter(hipStream_t s, lee_handle *h, const Block &b, double *y);
hipError_t plain_iblank_reset(hipStream_t s, const Block &b, double *y, int nv);
hipError_t aux_density(hipStream_t s, const Block &b, double gamma, const double *y, double *rho);
// HARMONIC_WAVE entropy factor e(t) = A (1 + sin(k c0 t)), k = 2 pi / wavelength
// (boundary_condition.cpp:232-344), and the round-jet Gaussian of a point.
double harmonic_entropy(const lee_config &c, double t);
double harmonic_gaussian(const lee_config &c, double x, double y, double z);
hipError_t apply_bcs(hipStream_t s, lee_handle *h, const Block &b, double *y64, float *y32, bool fp32,
                     double *rho64, float *rho32, double he = 0.0);
hipError_t overset_interpolate(hipStream_t s, lee_handle *h, bool fp32, int buf_index, bool last_stage);
hipError_t boundary_phase_fused(hipStream_t s, lee_handle *h, bool fp32, int tgt, bool last, double he);
// Overset interpolation of caller-owned FP64 block arrays ys[b] (device,
// 5n each), receiver values staged in buf (nrecv * NV doubles, device).
hipError_t overset_interpolate_arrays(hipStream_t s, lee_handle *h, double *const *ys, double *buf);
// Looks up the current device's static staging area of
// lee_update_boundary_noalloc (once per device; the device must be current).
hipError_t noalloc_prepare_device(int device);
hipError_t metrics2d(hipStream_t s, Block &b, const OpTable &fd);

hipError_t fused_stage(hipStream_t s, lee_handle *h, Block &b, int stage, const StageScalars &sc);
// 3-D (lee_kernels3d.hip)
hipError_t metrics3d(hipStream_t s, Block &b);
hipError_t deriv_xyz3(hipStream_t s, const Block &b, const double *f, double *out, int d);
hipError_t fused_stage3d(hipStream_t s, lee_handle *h, Block &b, int stage, const StageScalars &sc);
hipError_t fused_filter3d(hipStream_t s, lee_handle *h, Block &b, int src, int dst, int tmp);
hipError_t fused_filter(hipStream_t s, lee_handle *h, Block &b, int src, int dst);
// Two-block handles whose blocks share the kernel instance and forms run
// each stage, and the stacked filter, as one launch over both blocks.
bool pair_stage_ok(const lee_handle *h);
bool pair_filter_ok(const lee_handle *h);
hipError_t fused_stage_pair(hipStream_t s, lee_handle *h, int stage, const StageScalars &sc);
hipError_t fused_filter_pair(hipStream_t s, lee_handle *h, int src, int dst);
hipError_t batch_prepare(hipStream_t s, lee_handle *h);
hipError_t batch_stage(hipStream_t s, lee_handle *h, int stage);
hipError_t batch_filter(hipStream_t s, lee_handle *h);
void batch_release(lee_handle *h);
hipError_t adjoint_stage(hipStream_t s, lee_handle *h, Block &b, const AdjStage &st);
hipError_t adjoint_filter(hipStream_t s, lee_handle *h, Block &b, int src, int dst, bool with_mu);
// Transposed boundary update (overset interpolation, then the BC faces of
// every block in reverse order) on the adjoint buffer A[buf] of all blocks.
hipError_t adjoint_boundary(hipStream_t s, lee_handle *h, int buf);
hipError_t adjoint_zero(hipStream_t s, lee_handle *h, Block &b, int buf);  // buf < 0: the aux-density adjoint
hipError_t adjoint_zero_grad(hipStream_t s, Block &b);
hipError_t adjoint_add_rt_mu(hipStream_t s, lee_handle *h, Block &b, int buf);
hipError_t adjoint_stage3d(hipStream_t s, lee_handle *h, Block &b, const AdjStage &st);
hipError_t adjoint_filter3d(hipStream_t s, lee_handle *h, Block &b, int src, int dst, int scratch, bool with_mu);
hipError_t adjoint_add_rt_mu3(hipStream_t s, lee_handle *h, Block &b, int buf);
hipError_t observer_energy(hipStream_t s, lee_handle *h, Block &b, double *out);
hipError_t observer_gather(hipStream_t s, lee_handle *h, Block &b, double *out);
hipError_t observer_seed(hipStream_t s, lee_handle *h, Block &b, int buf, const double *samp, double scale);
hipError_t cast_state(hipStream_t s, const Block &b, int nvars, const double *src, float *dst, size_t n);
hipError_t cast_back(hipStream_t s, const Block &b, int nvars, const float *src, double *dst, size_t n);

}  // namespace lee
