/*
 * lee.h -- C ABI of camille_amd, the MI355X-native linearized-Euler (LEE)
 * time advance.  Plain pointers and sizes only; every call returns an int
 * status (0 = ok) and never exits the process.
 *
 * The reference has no plugin API or FFI: its seam is the C++ function
 *   void temporal::time_advance(UserInput*, StructuredGrid*, StructuredBlock*, State*)
 *   (src/solver/temporal_discretization.cpp:334), called once per step by
 *   solver::solve (src/solver/solver.cpp:41-89) after compute_CFL and
 *   `time_sol += dt`.
 * Each entry point below names the reference function it replaces.
 *
 * Arrays passed in or out are host memory in the reference layout: one
 * block = n = ni*nj*nk points, i fastest (Geometry::StructuredBlock::idx1D,
 * src/geometry/geometry.h:131-136); a field set is variable-major, e.g. a
 * solution is nv*n doubles ordered s', u', v', w', p' (src/core/param.h:64-71)
 * and then the passive scalars Z'_0 .. Z'_{ns-1} of a LEE_SCALAR run
 * (state.cpp:336-372), nv = 5 + lee_config.num_scalar.
 * The library copies what it needs; the caller keeps ownership.
 *
 * One handle = one set of blocks on the current HIP device, one stream;
 * calls on one handle must not overlap.
 */
#ifndef CAMILLE_AMD_LEE_H
#define CAMILLE_AMD_LEE_H

#ifdef __cplusplus
extern "C" {
#endif

#define LEE_OK 0
#define LEE_EINVAL 1      /* bad argument or unsupported option */
#define LEE_EHIP 2        /* HIP runtime error */
#define LEE_ENOMEM 3
#define LEE_ESTATE 4      /* call out of order (e.g. before lee_set_solution) */
#define LEE_EHOLES 5      /* too few fluid points between holes (spatial:1794) */

/* precision of the advance */
#define LEE_FP64 0        /* everything in FP64 */
#define LEE_FP32 1        /* FP32 state/stencils/RK; FP64 metrics, J, weights, time */

/* execution path */
#define LEE_PATH_FUSED 0  /* fused stage + filter kernels (default) */
#define LEE_PATH_PLAIN 1  /* one kernel per reference operator, FP64 only,
                             bit-identical arithmetic order */

/* boundary / buffer-zone types: src/core/param.h:104-125 numbering */
#define LEE_BC_DIRICHLET 0                   /* rejected: the reference exits for LEE (state.cpp:746-749) */
#define LEE_BC_DIRICHLET_ALLZERO 1
#define LEE_BC_DIRICHLET_HARMONICWAVE 2      /* the HARMONIC_WAVE state (boundary_condition.cpp:232-344) */
#define LEE_BC_NEUMANN 100
#define LEE_BC_WALL_SLIP_KINEMATIC 201
#define LEE_BC_CENTERLINE_CART_NORM2X 900
#define LEE_BC_CENTERLINE_CART_NORM2Y 901
#define LEE_BC_CENTERLINE_CART_NORM2Z 902
#define LEE_BC_CENTERLINE_AXISYM 903         /* the axis r = 0 of an axisymmetric run */
#define LEE_SPONGE_FREUND_AMBIENT 1000
#define LEE_SPONGE_FREUND_DIRICHLET 1001     /* rejected: the reference exits for LEE (state.cpp:746-749) */
#define LEE_SPONGE_FREUND_HARMONICWAVE 1002  /* relaxes to the HARMONIC_WAVE state (source.cpp:264-320) */

/* HARMONIC_WAVE settings (input.cpp:219-231): wave type and shape */
#define LEE_WAVE_NONE 0
#define LEE_WAVE_ACOUSTIC 1        /* refused like the reference does ("Unknown wave.", input.cpp:636-648) */
#define LEE_WAVE_ENTROPY 2
#define LEE_WAVEFORM_PLANE 1
#define LEE_WAVEFORM_HOMOGENEOUS 2
#define LEE_WAVEFORM_GAUSSIAN_ROUNDJET 3

/* physical sources (SIMULATION key; src/solver/source.cpp:22-107) */
#define LEE_SOURCE_NONE 0         /* CASE_TANNA_TPN49 and other source-free cases */
#define LEE_SOURCE_TWOCYLINDER 1  /* CASE_SCATTERING_TWOCYLINDER */
#define LEE_SOURCE_JET2D 2        /* CASE_2DJET_WITH_A_HARMONIC_SOURCE */
#define LEE_SOURCE_MONOPOLE 3     /* extension: A exp(-ln2 |x-x0|^2/w^2) sin(omega t);
                                     params {A, omega, x0, y0, z0, w} */
#define LEE_SOURCE_FIELD 4        /* extension: q(x) A sin(omega t) with a caller field q
                                     (lee_set_source_field); params {A, omega} */

typedef struct lee_handle lee_handle;

/* Scheme and run settings: the deck keys of src/core/input.cpp:99-284. */
typedef struct {
  int num_dim;             /* DIMENSION: 2, or 3 (fused path, sponges only, no overset) */
  double gamma;            /* GAMMA_SPECIFICHEAT */
  int fd_order;            /* SPATIAL_SCHEME STANDARD_CENTRAL ORDER_ACCURACY: 4;
                              11 = FDO11P (2-D; no batches, no adjoint) */
  int filter_order;        /* SPATIAL_FILTER STANDARD_CENTRAL ORDER_ACCURACY: 0 (NONE), 4, 8;
                              11 = SFO11P (2-D; no batches, no adjoint) */
  double filter_blend;     /* SPATIAL_FILTER BLEND */
  int rk_stages;           /* TEMPORAL_SCHEME: 2 (RUNGE_KUTTA2) or 4 (RUNGE_KUTTA4) */
  int fix_cfl;             /* FIX_DT_OR_CFL: 1 = FIX_CFL, 0 = FIX_DT */
  double cfl_or_dt;        /* CFL or DT */
  int sponge_order;        /* BUFFER_ZONE POLYNOMIAL_ORDER */
  double sponge_constant;  /* BUFFER_ZONE CONSTANT */
  int source_kind;         /* LEE_SOURCE_* */
  double source_params[8];
  int precision;           /* LEE_FP64 / LEE_FP32 */
  int path;                /* LEE_PATH_FUSED / LEE_PATH_PLAIN */
  double filter_strength;  /* SPATIAL_FILTER SFO11P STRENGTH (unused by the
                              standard filter, which runs at strength 1) */
  int axisym;              /* AXISYMMETRY TRUE (axis x, r = y; 2-D): the pressure
                              equation gains -gamma (pbar u_r' + p' ubar_r) / r */
  int harmonic_wave;       /* HARMONIC_WAVE: LEE_WAVE_*; the state that
                              dirichlet_harmonicwave faces impose and
                              sponge_freund_harmonicwave zones relax to */
  int harmonic_shape;      /* HARMONIC_WAVE SHAPE: LEE_WAVEFORM_* */
  int harmonic_direction;  /* HARMONIC_WAVE DIRECTION: 0 x, 1 y, 2 z */
  double harmonic_amplitude;   /* AMPLITUDE, in [0, 1] */
  double harmonic_wavelength;  /* WAVELENGTH */
  double harmonic_halfwidth;   /* HALFWIDTH (WAVEFORM_GAUSSIAN_ROUNDJET) */
  int num_scalar;          /* PHYSICAL_MODEL LEE_SCALAR NUMBER_SCALAR (input.cpp:113-119), 0 to
                              LEE_MAX_SCALAR; 0 = LEE.  Scalars Z'_k: dZ'/dt = -ubar . grad Z'
                              - u' . grad Zbar (spatial:2134-2158), filtered, relaxed to 0 in
                              sponges, Neumann at wall-slip and centerline faces (2-D single
                              handles; no batches, no adjoint) */
} lee_config;

#define LEE_MAX_SCALAR 4

/* One BC-file region (src/io/io.cpp:1158-1268) with negative indices
 * resolved: type, normal direction (0 xi, 1 eta, 2 zeta), end (0 = '+' low
 * face, 1 = '-' high face) and the block-index box is..ie (inclusive). */
typedef struct {
  int type, dir, end;
  int is[3], ie[3];
} lee_boundary;

/* One block: grid coordinates (3n: x, y, z) and base state (nv*n: s, u, v,
 * w, p and the scalars' means) as read from GRID_FILE and BASESTATE_FILE;
 * nk = 1 in 2-D. */
typedef struct {
  int ni, nj, nk;
  const double *xyz;
  const double *mean;
  int nboundary;
  const lee_boundary *boundary;
} lee_block;

/* Explicit overset interpolation tables (the Overture file of
 * src/io/overture.cpp:473-680, flattened): per receiver its block and 1-D
 * index, its donor block and the 1-D index of the stencil's lower-left
 * corner, and width[0]*width[1]*width[2] raw coefficients (w1 fastest) of
 * the donor block's stencil; holes as (block, 1-D index). */
typedef struct {
  int nrecv;
  const int *recv_block, *recv_l;
  const int *donor_block, *donor_l;
  const int *widths;       /* 3 per block */
  const double *coef;      /* nrecv * coef_stride */
  int coef_stride;
  int nhole;
  const int *hole_block, *hole_l;
} lee_overset;

/* simulation::initialize from the base state onwards (src/core/simulation.cpp:
 * 164-240): uploads the blocks, computes metrics and Jacobians (before any
 * hole, metrics.cpp:9-44), marks holes and receivers, computes sponge
 * strengths (geometry.cpp:1028-1142) and mean gradients (spatial.cpp:2250-2271),
 * and prepares the CFL evaluation.  `ov` may be NULL. */
int lee_create(const lee_config *cfg, const lee_block *blocks, int nblocks, const lee_overset *ov,
               lee_handle **out);
void lee_destroy(lee_handle *h);

/* Message of the last failed call on this handle (or of lee_create when h is NULL). */
const char *lee_last_error(const lee_handle *h);

/* Restart / IC: io::read_solution (plot3d.cpp:715-833) + compute_dependent_variables.
 * sol = nv*n doubles; sets time_step and time_sol like tau[2], tau[3]. */
int lee_set_solution(lee_handle *h, int iblock, const double *sol);
int lee_get_solution(lee_handle *h, int iblock, double *sol);
int lee_set_time(lee_handle *h, int time_step, double time_sol, double dt);
int lee_get_time(const lee_handle *h, int *time_step, double *time_sol, double *dt);

/* temporal::compute_CFL (temporal_discretization.cpp:397-450) at the
 * current dt, with the reference's integer truncation of |U m|. */
int lee_compute_cfl(lee_handle *h, double *cfl_max);

/* temporal::time_advance (temporal_discretization.cpp:334-376): one step
 * from the current solution with the handle's dt, sources and boundary
 * data evaluated at time_sol + c_i dt (time_sol already advanced). */
int lee_time_advance(lee_handle *h);

/* solver::solve (solver.cpp:41-89) for nsteps: per step compute_CFL, dt
 * rescaled under FIX_CFL, time_sol += dt, time_advance.  Enqueued
 * asynchronously; lee_synchronize waits. */
int lee_solve(lee_handle *h, int nsteps);
int lee_synchronize(lee_handle *h);

/* Unit entry points on one block (host arrays, nv*n doubles).
 * spatial::compute_RHS (spatial:1963-1990) of y at `time`, with the aux
 * density taken from y itself. */
int lee_compute_rhs(lee_handle *h, int iblock, const double *y, double time, double *rhs);
/* spatial::apply_filter (spatial:2191-2246), in place. */
int lee_apply_filter(lee_handle *h, int iblock, double *y);
/* StandardCentral::take_derivative_xyz (spatial:1802-1826) of one field. */
int lee_take_derivative_xyz(lee_handle *h, int iblock, const double *f, int dir_xyz, double *out);
/* spatial::update_boundary (spatial:2162-2187): BCs then overset
 * interpolation, in place on all blocks (ys[b] = nv*n doubles of block b);
 * time-dependent boundary data (a harmonic wave) at the handle's time_sol. */
int lee_update_boundary(lee_handle *h, double *const *ys);

/* lee_update_boundary without any allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).
 * ys_dev[b] = nv*n doubles of block b in DEVICE memory, updated in place:
 * BCs, then overset interpolation, enqueued on the handle's stream
 * (lee_synchronize waits; the caller orders its own streams around it).
 * The receiver values are staged in a static device array of
 * LEE_NOALLOC_MAX_RECV * 5 doubles per device (about 5 MiB), held until the
 * process ends; a handle whose receivers need more (nrecv * nv doubles)
 * gets LEE_EINVAL.  All
 * calls on one device share that array: calls that overlap in time on one
 * device race.  No hipMalloc, hipFree or host copy is made. */
#define LEE_NOALLOC_MAX_RECV 131072
int lee_update_boundary_noalloc(lee_handle *h, double *const *ys_dev);

/* Init products, for checks: metrics[l][i] (9n), J (n), iblank (n),
 * mean gradients (nv*num_dim*n, ivar*num_dim+dir), summed sponge strength (n). */
int lee_get_metrics(lee_handle *h, int iblock, double *metrics, double *jac);
int lee_get_iblank(lee_handle *h, int iblock, int *iblank);
int lee_get_mean_gradient(lee_handle *h, int iblock, double *grad);
int lee_get_sponge(lee_handle *h, int iblock, double *sigma);

/* Timing of the last lee_solve on the device stream (HIP events) and the
 * number of stage-kernel launches it made; for benchmarks. */
int lee_last_solve_ms(lee_handle *h, float *ms);

/* Per-kernel timing of the fused path for benchmarks (no reference
 * counterpart): while enabled, HIP events bracket every launch on the
 * handle's stream.  lee_get_kernel_timing sums the recorded durations since
 * the last call into ms[5] / count[5] by class: 0 first RK stage, 1 middle
 * stages, 2 last stage, 3 filter, 4 boundary conditions + overset. */
int lee_set_kernel_timing(lee_handle *h, int on);
int lee_get_kernel_timing(lee_handle *h, double *ms, int *count);

/* Tile shapes of the fused kernels (no reference counterpart).  The first
 * lee_solve of a handle times every stage tile height (4 * stage_rows, with
 * stage_rows 1, 2 or 4) and filter tile height (4, 8, 12 or 16) on each block
 * and keeps the fastest, cached per (ni, nj, precision, mean kind, holes).
 * lee_autotune re-runs the search; lee_set_tile_shape pins a shape (every
 * shape gives bit-identical results). */
int lee_autotune(lee_handle *h);
int lee_set_tile_shape(lee_handle *h, int iblock, int stage_rows, int filter_rows);
int lee_get_tile_shape(lee_handle *h, int iblock, int *stage_rows, int *filter_rows);
/* Launch form of the fused filter (no reference counterpart): workgroups 0
 * launches one workgroup per tile; a positive count (rounded up to a
 * multiple of 8) launches that many persistent workgroups, each loading its
 * next tile while it filters the current one.  The autotuner also tries 2, 4
 * and 8 workgroups per CU.  Results are bit-identical in every form.  Setting
 * the tile shape or the grid selects the tile forms. */
int lee_set_filter_grid(lee_handle *h, int iblock, int workgroups);
int lee_get_filter_grid(lee_handle *h, int iblock, int *workgroups);
/* Column-march forms of the fused filter (no reference counterpart).  A
 * wave filters one variable on a strip of 64 grid columns (56 of them
 * output), taking the xi neighbours from adjacent lanes and keeping the eta
 * window in registers.  waves 1: each wave filters `rows` rows alone (8 or
 * 16; blocks without holes); waves 4 or 8: a workgroup stacks that many
 * waves of `rows` rows each (4, 8 or 16 for 4 waves; 8 for 8) and shares
 * their xi rows through LDS (blocks with holes too).  (0, 0) returns to the
 * tile forms above.  A batch (one launch for every solve) and the 3-D plane
 * pass take stacks of 4 waves only.  The autotuners try these forms;
 * results are bit-identical to the tile forms. */
int lee_set_filter_march(lee_handle *h, int iblock, int waves, int rows);
int lee_get_filter_march(lee_handle *h, int iblock, int *waves, int *rows);

/* Batched ensemble (no reference counterpart: the reference runs one
 * problem per process).  nsolve independent single-block problems, block k
 * with its own lee_config cfgs[k], advanced together: per RK stage one
 * persistent launch covers the tiles of every solve, and one step is
 * captured once in a HIP graph and replayed.  The scheme settings (num_dim,
 * gamma, orders, blend, stages, sponge law, precision, path) must agree;
 * source, CFL or dt, grid, sponges and mean flow may differ (the means must
 * be all uniform or all non-uniform).  Fused path only, filter required,
 * no overset.  Each solve keeps its own clock (FIX_CFL scales its own dt);
 * the per-block entry points take the solve index as `iblock`, lee_solve
 * advances every solve by nsteps, lee_get_time reports solve 0. */
int lee_batch_create(const lee_config *cfgs, const lee_block *blocks, int nsolve, lee_handle **out);
int lee_get_solve_time(const lee_handle *h, int isolve, int *time_step, double *time_sol, double *dt);
/* Launch settings of a batch: graph replay on/off, persistent grid sizes of
 * the stage and filter launches (0 = one workgroup per tile;
 * LEE_GRID_RESIDENT = as many workgroups as the device holds at once for
 * that kernel, its occupancy per CU times the CUs).  Chosen by the
 * autotuner at the first lee_solve unless set here; every setting gives
 * bit-identical results. */
#define LEE_GRID_RESIDENT (-1)
int lee_set_batch_launch(lee_handle *h, int use_graph, int grid_stage, int grid_filter);
int lee_get_batch_launch(lee_handle *h, int *use_graph, int *grid_stage, int *grid_filter);

/* Discrete adjoint (no reference counterpart: the reference has no
 * adjoint).  Fused path; 2-D blocks with sponges (harmonic-wave zones
 * included), sources, holes, overset interpolation, the axisymmetric term,
 * passive scalars (they feed nothing back: the adjoint runs on the five
 * LEE variables) and the dirichlet_allzero, harmonic-wave, Neumann,
 * wall-slip, Cartesian and axisymmetric centerline faces; 3-D blocks with
 * sponges and sources.  Not covered: the 11-point schemes.
 *
 * Source field of a LEE_SOURCE_FIELD block: p-equation source q(x) A sin(omega t). */
int lee_set_source_field(lee_handle *h, int iblock, const double *q);
/* Observer functional J = sum over steps k in (window_start, nsteps] of
 * sum_l w_l p'_k(l)^2, p'_k the pressure after step k at the observer
 * points (block[l], index[l]); nobs = 0 clears them. */
int lee_set_observer(lee_handle *h, int nobs, const int *block, const int *index, const double *weight);
/* p' of the current state at the observer points, in the order given to
 * lee_set_observer (nobs doubles); for probe time series. */
int lee_sample_observers(lee_handle *h, double *out);
/* Runs nsteps steps from the current state and clock (like lee_solve),
 * keeping the state every ckpt_every steps on the device, and returns J. */
int lee_objective_forward(lee_handle *h, int nsteps, int window_start, int ckpt_every, double *J);
/* Reverse pass of the last lee_objective_forward for the gradient of
 * scale * J: each checkpoint segment is recomputed forwards (observer
 * samples kept), then the transposed filter, stages and RK updates run
 * backwards through it.  Afterwards the state and clock are those at the
 * end of the forward run again. */
int lee_objective_backward(lee_handle *h, double scale);
/* dJ/dq of a LEE_SOURCE_FIELD block (n doubles) and dJ/dy0 (5n doubles,
 * the state the forward run started from). */
int lee_get_source_gradient(lee_handle *h, int iblock, double *grad_q);
int lee_get_state_gradient(lee_handle *h, int iblock, double *grad_y0);
/* Transposed unit operators, for dot-product checks against
 * lee_compute_rhs (without its source term) and lee_apply_filter:
 * out = L^T kappa with the aux density of the argument, and y <- F^T y. */
int lee_compute_rhs_adjoint(lee_handle *h, int iblock, const double *kappa, double *out);
/* Transposed boundary update, for dot-product checks against
 * lee_update_boundary: ys <- B^T ys on every block (overset interpolation
 * transposed, then each block's BC faces in reverse order). */
int lee_update_boundary_adjoint(lee_handle *h, double *const *ys);
int lee_apply_filter_adjoint(lee_handle *h, int iblock, double *y);

/* Library identification: returns a static string (build info). */
const char *lee_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CAMILLE_AMD_LEE_H */
