/*
 * lee_oracle.c -- CPU restatement of camille's linearized-Euler (LEE) time
 * advance, in FP64, following the reference loop by loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (camille_amd/) links,
 * loads or calls this file.  It is the checker used by tests/, by the smoke
 * entry point and by the CPU-baseline leg of bench.py.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference's src/).  Arithmetic is written in the same operation
 * order as the reference so that, compiled with -ffp-contract=off, it
 * reproduces the reference binary bit for bit (checked against the golden
 * fixtures written by tests/golden/make_golden.py).
 *
 * Parity pins: FD, filter, metrics, sponge, CFL, RK, aux, source time,
 * wall-slip/Neumann BCs and hole-free paths are pinned against reference
 * fixtures; hole closures and overset interpolation are pinned by closed
 * forms in tests/test_oracle.py (polynomial exactness, normalisation).
 *
 * Layout: one block = SoA double arrays of n = ni*nj*nk points, i fastest,
 * exactly the reference's single-rank grid (no ghost cells, is = iso = 0,
 * ie = ieo = n_dir - 1; src/geometry/geometry.cpp:81-212).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXB 8
#define NV 5   /* s', u'_x, u'_y, u'_z, p' */
#define MAXV 9 /* NV + up to 4 passive scalars Z'_k (LEE_SCALAR, input.cpp:768-774) */
#define IS_ 0 /* s'  (param.h IVAR_S)  */
#define IUX 1
#define IUY 2 /* u_y, also IVAR_UR of an axisymmetric run (param.h:74) */
#define IP 4
#define NOTBLANKED 0
#define BLANKED 1
#define INTERPOLATED (-1)

/* boundary type codes of src/core/param.h:104-125 */
#define BC_DIRICHLET_ALLZERO 1
#define BC_NEUMANN 100
#define BC_WALL_SLIP 201
#define BC_CL_X 900
#define BC_CL_Z 902
#define BC_DIRICHLET_HARMONIC 2
#define SPONGE_AMBIENT 1000
#define SPONGE_HARMONIC 1002
#define WAVE_ENTROPY 2
#define WAVEFORM_HOMOGENEOUS 2
#define WAVEFORM_GAUSSIAN_ROUNDJET 3

#define SRC_NONE 0
#define SRC_TWOCYL 1 /* CASE_SCATTERING_TWOCYLINDER, source.cpp:60-92 */
#define SRC_JET 2    /* CASE_2DJET_WITH_A_HARMONIC_SOURCE, source.cpp:27-58 */
#define SRC_MONOPOLE 3 /* time-harmonic Gaussian monopole (not in the reference) */

static char g_err[512];
/* Deliberate-mistake switch used only by the oracle's own self-tests, to
 * show that the golden fixtures catch each quirk (0 = faithful):
 *  1 fabs instead of int abs in the CFL      2 sources at t_n instead of t_{n+1}
 *  3 aux density refreshed after the filter  4 metrics transposed in d/dx_i
 *  5 mean-gradient terms dropped             6 filter sweeps eta before xi
 *  7 right hole closures after any non-hole cell (`!=` for `/=`)
 *  8 monopole with cos(omega t) for sin(omega t)
 *  9 monopole Gaussian exp(-|x-x0|^2/w^2), the ln 2 of the half-width dropped
 * 10 entropy wave with cos(k c0 t) for sin(k c0 t)
 * 11 passive scalars left out of the filter
 * 12 wall-slip leaving the scalars out of its Neumann update */
static int g_plant = 0;
void orc_set_plant(int code) { g_plant = code; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char *orc_error(void) { return g_err; }

typedef struct {
  int half;          /* half_size_of_stencil */
  double st[11];     /* interior stencil, index s + half */
  int nb, sb;        /* num_of_boundaryCells, size_of_boundaryStencil */
  double bnd[5][9];  /* stencil_boundary[ib][s] */
  double sign_right; /* multi_fac_4boundaryRight */
  int ghost;         /* num_cells_ghost_{finiteDifference,filter} */
} op_t;

typedef struct {
  int type, dir, end;   /* which_model, which_dir, which_end */
  int is[3], ie[3];     /* zone/face box in block indices */
  double *sigma;        /* sponge strength over the box (sponges only) */
} bnd_t;

typedef struct {
  int ni, nj, nk, n, nd[3];
  double *xyz[3];
  double *met;   /* [9][n] metrics[l][i] at met[(3l+i)n + p] */
  double *minv;  /* [9][n] metricsInverse[i][l] */
  double *jac;
  int *iblank;
  double *mean[MAXV], *mgrad[MAXV * 3], *aux[4];
  double *sol[MAXV], *sol_old[MAXV];
  double *rhs[MAXV], *interim[MAXV], *yn[MAXV], *ynp1[MAXV];
  int nbnd;
  bnd_t bnd[32];
  /* operator extents per op: [0] derivative, [1] filter */
  int is4[2][3], ie4[2][3], iso4[2][3], ieo4[2][3];
} blk_t;

typedef struct {
  int nrecv;
  int *rb, *rl;        /* receiver block, receiver 1-D index */
  int *db, *dl;        /* donor block, donor lower-left 1-D index */
  int w[MAXB][3];      /* stencil widths of the donor block */
  double *coef;        /* nrecv * maxstencil normalised coefficients */
  int maxst;
  double *buf;         /* nrecv * nv */
} ovs_t;

typedef struct {
  int nblk, ndim;
  int nv;             /* NV + number of passive scalars */
  double gamma;
  op_t fd, flt;
  int do_filter;
  double filter_blend;
  int rk_stages;
  double b_i[4], c_i[4], a_ij[4][4];
  int sponge_order;
  double sponge_constant;
  int fix_cfl;
  double cfl_target, dt, time_sol;
  int time_step;
  int src_kind;
  double src[8];
  int do_overset;
  int axisym;
  int hw_wave, hw_shape, hw_dir;      /* HARMONIC_WAVE (input.cpp:219-231) */
  double hw_amp, hw_wavelength, hw_halfwidth;
  ovs_t ov;
  double *dflux, *dflux_sum, *flux, *func0, *func; /* scratch of max n */
  int nmax;
  blk_t b[MAXB];
  double pi, twopi, inv_twopi;
} orc_t;

/* ---------------------------------------------------------------- tables */

/* OptimizedCentralExplicit::initialize, FDO11P (spatial_discretization.cpp:
 * 228-344): Bogey-Bailly FDo11p interior; closure rows i=0 and i=1 Berland's
 * FD06 and FD15 (7 points from the boundary), i=2 the standard 5-point
 * centred row, i=3 FDo7p, i=4 FDo9p, every row anchored at the boundary;
 * 9 entries per row, mirrored with a sign flip on the right. */
static void init_fdo11p(op_t *op) {
  op->half = 5;
  op->nb = 5;
  op->sb = 9;
  op->sign_right = -1.0;
  op->ghost = 5; /* input.cpp:698-699 */
  double *st = op->st + op->half;
  st[0] = 0.0; st[1] = 0.872756993962; st[2] = -0.286511173973; st[3] = 0.090320001280;
  st[4] = -0.020779405824; st[5] = 0.002484594688;
  const double r0[7] = {-2.225833963270, 4.827779580575, -5.001388453836, 3.911103941646, -2.115267458633,
                        0.718882784412, -0.115276430894};
  const double r1[7] = {-0.212932721951, -1.060320390770, 2.078926116439, -1.287179452384, 0.685176395471,
                        -0.245320613994, 0.041650667189};
  const double r2[5] = {1.0 / 12.0, -8.0 / 12.0, 0.0, 8.0 / 12.0, -1.0 / 12.0};
  const double r3[7] = {-0.02625099337517627, 0.18833730683403838, -0.797921633542548, 0.0, 0.797921633542548,
                        -0.18833730683403838, 0.02625099337517627};
  const double r4[9] = {0.007650904064, -0.059463584768, 0.244678631765, -0.841570125482, 0.0,
                        0.841570125482, -0.244678631765, 0.059463584768, -0.007650904064};
  for (int k = 0; k < 7; k++) { op->bnd[0][k] = r0[k]; op->bnd[1][k] = r1[k]; op->bnd[3][k] = r3[k]; }
  for (int k = 0; k < 5; k++) op->bnd[2][k] = r2[k];
  for (int k = 0; k < 9; k++) op->bnd[4][k] = r4[k];
  for (int s = 1; s <= op->half; s++) st[-s] = -st[s];
}

/* OptimizedCentralFilterExplicit::initialize, SFO11P (spatial_discretization.
 * cpp:348-492): Bogey-Bailly SFo11p interior at strength sigma (1 - sigma d0
 * at the centre, -sigma d_k off it); closure rows 0 and 1 identity, 2 and 3
 * Lele's 4th- and 6th-order filters (at full strength), row 4 the 9-point
 * optimized filter at strength sigma; mirrored on the right. */
static void init_sfo11p(op_t *op, double strength) {
  op->half = 5;
  op->nb = 5;
  op->sb = 9;
  op->sign_right = 1.0;
  op->ghost = 5; /* input.cpp:718-719 */
  double *st = op->st + op->half;
  st[0] = 0.215044884112; st[1] = -0.187772883589; st[2] = 0.123755948787; st[3] = -0.059227575576;
  st[4] = 0.018721609157; st[5] = -0.002999540835;
  st[0] = 1.0 - strength * st[0];
  for (int s = 1; s <= op->half; s++) st[s] *= -strength;
  op->bnd[0][0] = 1.0;
  op->bnd[1][1] = 1.0;
  op->bnd[2][0] = (-1.0 / 8.0) / 2.0; op->bnd[2][1] = (1.0 / 2.0) / 2.0; op->bnd[2][2] = 5.0 / 8.0;
  op->bnd[2][3] = (1.0 / 2.0) / 2.0; op->bnd[2][4] = (-1.0 / 8.0) / 2.0;
  op->bnd[3][0] = (1.0 / 32.0) / 2.0; op->bnd[3][1] = (-3.0 / 16.0) / 2.0; op->bnd[3][2] = (15.0 / 32.0) / 2.0;
  op->bnd[3][3] = 11.0 / 16.0; op->bnd[3][4] = (15.0 / 32.0) / 2.0; op->bnd[3][5] = (-3.0 / 16.0) / 2.0;
  op->bnd[3][6] = (1.0 / 32.0) / 2.0;
  const double r4[9] = {0.008228661760, -0.045211119360, 0.120007591680, -0.204788880640, 0.243527493120,
                        -0.204788880640, 0.120007591680, -0.045211119360, 0.008228661760};
  for (int k = 0; k < 9; k++) op->bnd[4][k] = k == 4 ? 1.0 - strength * r4[k] : r4[k] * -strength;
  for (int s = 1; s <= op->half; s++) st[-s] = st[s];
}

/* StandardCentral::initialize, solver/spatial_discretization.cpp:28-109;
 * order 11 selects FDO11P */
static int init_fd(op_t *op, int order) {
  memset(op, 0, sizeof *op);
  if (order == 11) { init_fdo11p(op); return 0; }
  /* order 2 exists in the reference in name only: its one-entry closure
   * table is written at index 1 (spatial:52,70-71) and runs diverge */
  if (order != 4) { set_err("FD order must be 4 (the reference's order-2 closure is broken)"); return -1; }
  op->half = order / 2;
  op->nb = op->half;
  op->sb = (order - 2) + 1;
  op->sign_right = -1.0;
  op->ghost = order / 2; /* input.cpp:694-707 */
  double *st = op->st + op->half;
  if (order == 2) {
    st[0] = 0.0; st[1] = 1.0 / 2.0;
    op->bnd[0][0] = -1.0; op->bnd[0][1] = 1.0;
  } else {
    st[0] = 0.0; st[1] = 8.0 / 12.0; st[2] = -1.0 / 12.0;
    op->bnd[0][0] = -3.0 / 2.0; op->bnd[0][1] = 4.0 / 2.0; op->bnd[0][2] = -1.0 / 2.0;
    op->bnd[1][0] = -1.0 / 2.0; op->bnd[1][1] = 0.0; op->bnd[1][2] = 1.0 / 2.0;
  }
  for (int s = 1; s <= op->half; s++) st[-s] = -st[s];
  return 0;
}

/* StandardCentralFilter::initialize, spatial_discretization.cpp:113-224;
 * order 11 selects SFO11P at the given strength */
static int init_filter(op_t *op, int order, double strength) {
  memset(op, 0, sizeof *op);
  if (order == 11) { init_sfo11p(op, strength); return 0; }
  if (order != 4 && order != 8) { set_err("filter order must be 4 or 8"); return -1; }
  op->half = order / 2;
  op->nb = op->half;
  op->sb = (order - 2) + 1;
  op->sign_right = 1.0;
  op->ghost = order / 2; /* input.cpp:711-727 */
  double *st = op->st + op->half;
  if (order == 4) {
    st[0] = 5.0 / 8.0; st[1] = (1.0 / 2.0) / 2.0; st[2] = (-1.0 / 8.0) / 2.0;
    op->bnd[0][0] = 1.0;
    op->bnd[1][1] = 1.0;
  } else {
    st[0] = 35.0 / 128.0; st[1] = -7.0 / 32.0; st[2] = 7.0 / 64.0; st[3] = -1.0 / 32.0; st[4] = 1.0 / 256.0;
    st[0] = 1.0 - strength * st[0];
    for (int s = 1; s <= op->half; s++) st[s] *= -strength;
    op->bnd[0][0] = 1.0;
    op->bnd[1][1] = 1.0;
    op->bnd[2][0] = (-1.0 / 8.0) / 2.0; op->bnd[2][1] = (1.0 / 2.0) / 2.0; op->bnd[2][2] = 5.0 / 8.0;
    op->bnd[2][3] = (1.0 / 2.0) / 2.0; op->bnd[2][4] = (-1.0 / 8.0) / 2.0;
    op->bnd[3][0] = (1.0 / 32.0) / 2.0; op->bnd[3][1] = (-3.0 / 16.0) / 2.0; op->bnd[3][2] = (15.0 / 32.0) / 2.0;
    op->bnd[3][3] = 11.0 / 16.0; op->bnd[3][4] = (15.0 / 32.0) / 2.0; op->bnd[3][5] = (-3.0 / 16.0) / 2.0;
    op->bnd[3][6] = (1.0 / 32.0) / 2.0;
  }
  for (int s = 1; s <= op->half; s++) st[-s] = st[s];
  return 0;
}

static inline int idx(const blk_t *b, int i, int j, int k) { return b->ni * b->nj * k + b->ni * j + i; }
static inline int stride_of(const blk_t *b, int dir) { return dir == 0 ? 1 : (dir == 1 ? b->ni : b->ni * b->nj); }

/* override_local_indices_4derivative/4filter, geometry.cpp:931-1001, for a
 * grid that owns its whole block (no internal neighbours). */
static void set_extents(orc_t *o, blk_t *b) {
  for (int which = 0; which < 2; which++) {
    int g = which == 0 ? o->fd.ghost : o->flt.ghost;
    for (int d = 0; d < 3; d++) {
      b->is4[which][d] = 0;
      b->ie4[which][d] = b->nd[d] - 1;
      b->iso4[which][d] = 0;
      b->ieo4[which][d] = b->nd[d] - 1;
      if (b->nd[d] > 1) { b->is4[which][d] += g; b->ie4[which][d] -= g; }
    }
    for (int d = 0; d < o->ndim; d++) {
      b->iso4[which][d] = b->is4[which][d] - g;
      b->ieo4[which][d] = b->ie4[which][d] + g;
    }
  }
}

/* ------------------------------------------------------- hole corrections */

/* check_boundary_stencils_if_all_fluid_points, spatial:1749-1798 */
static int check_fluid(const blk_t *b, const op_t *op, const int *ijk, int dir, int left) {
  int sum = 0;
  for (int ib = 0; ib < op->nb; ib++) {
    int w[3] = {ijk[0], ijk[1], ijk[2]};
    w[dir] += left ? ib : -1 - ib;
    if (w[dir] < 0 || w[dir] >= b->nd[dir]) { set_err("hole closure leaves the block"); return -1; }
    int v = b->iblank[idx(b, w[0], w[1], w[2])];
    sum += v > 0 ? v : 0;
  }
  if (sum != NOTBLANKED) { set_err("too few fluid points between holes"); return -1; }
  return 0;
}

/* insert_right_boundary_stencils, spatial:1656-1700 */
static void insert_right(const blk_t *b, const op_t *op, const int *ijk, const double *f, double *df, int dir) {
  int st = stride_of(b, dir);
  int w[3] = {ijk[0], ijk[1], ijk[2]};
  w[dir] -= 1;
  int lref = idx(b, w[0], w[1], w[2]);
  for (int ib = 0; ib < op->nb; ib++) {
    int v[3] = {ijk[0], ijk[1], ijk[2]};
    v[dir] += -1 - ib;
    if ((v[dir] - 0) * (v[dir] - (b->nd[dir] - 1)) <= 0) {
      int l1 = idx(b, v[0], v[1], v[2]);
      df[l1] = 0.0;
      for (int s = 0; s < op->sb; s++)
        df[l1] += op->sign_right * op->bnd[ib][s] * f[lref - s * st];
    }
  }
}

/* insert_left_boundary_stencils, spatial:1704-1745 */
static void insert_left(const blk_t *b, const op_t *op, const int *ijk, const double *f, double *df, int dir) {
  int st = stride_of(b, dir);
  int lref = idx(b, ijk[0], ijk[1], ijk[2]);
  for (int ib = 0; ib < op->nb; ib++) {
    int v[3] = {ijk[0], ijk[1], ijk[2]};
    v[dir] += ib;
    if ((v[dir] - 0) * (v[dir] - (b->nd[dir] - 1)) <= 0) {
      int l1 = idx(b, v[0], v[1], v[2]);
      df[l1] = 0.0;
      for (int s = 0; s < op->sb; s++)
        df[l1] += op->bnd[ib][s] * f[lref + s * st];
    }
  }
}

/* StandardCentral::iblank_correction, spatial:1356-1652.  The hole-after-
 * fluid branch tests `last_iblank /= BLANKED`, i.e. last_iblank != 0: right
 * closures are inserted only after an INTERPOLATED cell (:1391). */
static int iblank_correction(const blk_t *b, const op_t *op, int which, const double *f, double *df, int dir) {
  const int *iso = b->iso4[which], *ieo = b->ieo4[which];
  int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
  for (int c2 = iso[o2]; c2 <= ieo[o2]; c2++)
    for (int c1 = iso[o1]; c1 <= ieo[o1]; c1++) {
      int ijk[3];
      ijk[o1] = c1; ijk[o2] = c2; ijk[dir] = iso[dir];
      int last = b->iblank[idx(b, ijk[0], ijk[1], ijk[2])];
      for (int i = iso[dir]; i <= ieo[dir]; i++) {
        ijk[dir] = i;
        int l0 = idx(b, ijk[0], ijk[1], ijk[2]);
        int my = b->iblank[l0];
        if (my == BLANKED) {
          if (last == BLANKED) {
            df[l0] = 0.0;
          } else if (g_plant == 7 ? last != BLANKED : last != 0) { /* the reference's `last_iblank /= BLANKED` */
            if (i <= 0) {
            } else {
              if (check_fluid(b, op, ijk, dir, 0)) return -1;
              insert_right(b, op, ijk, f, df, dir);
            }
          }
        } else {
          if (last == BLANKED) {
            if (i > b->nd[dir] - 1) {
            } else {
              if (check_fluid(b, op, ijk, dir, 1)) return -1;
              insert_left(b, op, ijk, f, df, dir);
            }
          }
        }
        last = my;
      }
    }
  return 0;
}

/* StandardCentral::apply_spatial_discretization, spatial:1078-1290.
 * `which` selects the derivative (0) or filter (1) extents. */
static int apply_op(const blk_t *b, const op_t *op, int which, const double *f, double *df, int dir) {
  if (b->nd[dir] == 1) return 0;
  for (int l = 0; l < b->n; l++) df[l] = 0.0;
  const int *is = b->is4[which], *ie = b->ie4[which], *iso = b->iso4[which], *ieo = b->ieo4[which];
  int st = stride_of(b, dir);
  const double *c = op->st + op->half;
  int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
  for (int c2 = iso[o2]; c2 <= ieo[o2]; c2++)
    for (int c1 = iso[o1]; c1 <= ieo[o1]; c1++) {
      int ijk[3];
      ijk[o1] = c1; ijk[o2] = c2;
      /* interior */
      for (int i = is[dir]; i <= ie[dir]; i++) {
        ijk[dir] = i;
        int l0 = idx(b, ijk[0], ijk[1], ijk[2]);
        for (int s = -op->half; s <= op->half; s++) df[l0] += c[s] * f[l0 + s * st];
      }
      /* left closure, anchored at iso */
      ijk[dir] = iso[dir];
      int lso = idx(b, ijk[0], ijk[1], ijk[2]);
      for (int i = iso[dir]; i < is[dir]; i++) {
        int il = i - iso[dir];
        ijk[dir] = i;
        int l0 = idx(b, ijk[0], ijk[1], ijk[2]);
        for (int s = 0; s < op->sb; s++) df[l0] += op->bnd[il][s] * f[lso + s * st];
      }
      /* right closure, anchored at ieo, mirrored */
      ijk[dir] = ieo[dir];
      int leo = idx(b, ijk[0], ijk[1], ijk[2]);
      for (int i = ieo[dir]; i > ie[dir]; i--) {
        int il = ieo[dir] - i;
        ijk[dir] = i;
        int l0 = idx(b, ijk[0], ijk[1], ijk[2]);
        for (int s = 0; s < op->sb; s++) df[l0] += op->sign_right * op->bnd[il][s] * f[leo - s * st];
      }
    }
  return iblank_correction(b, op, which, f, df, dir);
}

/* take_derivative_xi_eta_zeta, spatial:1322-1336 */
static int deriv_xi(orc_t *o, blk_t *b, const double *f, int dir) {
  for (int l = 0; l < b->n; l++) o->dflux[l] = 0.0;
  return apply_op(b, &o->fd, 0, f, o->dflux, dir);
}

/* take_derivative_xyz, spatial:1802-1826: sum_l metrics[l][i] d/dxi_l, then * J */
static int deriv_xyz(orc_t *o, blk_t *b, const double *f, int dir_x) {
  double *sum = o->dflux_sum;
  for (int l = 0; l < b->n; l++) sum[l] = 0.0;
  for (int l = 0; l < o->ndim; l++) {
    if (deriv_xi(o, b, f, l)) return -1;
    const double *m = b->met + (size_t)(g_plant == 4 ? 3 * dir_x + l : 3 * l + dir_x) * b->n;
    for (int p = 0; p < b->n; p++) sum[p] += m[p] * o->dflux[p];
  }
  for (int p = 0; p < b->n; p++) o->dflux[p] = sum[p] * b->jac[p];
  return 0;
}

/* ------------------------------------------------------------ init pieces */

/* metrics::compute (ThomasLombard 2-D), metrics.cpp:9-129 + Jacobian
 * :255-328 + inverse :355-426.  3-D has no Jacobian in the reference (it
 * exits, :290-292); the oracle rejects 3-D grids the same way. */
static int compute_metrics(orc_t *o, blk_t *b) {
  int n = b->n;
  if (o->ndim != 2) { set_err("3-D Jacobian computation is not implemented (reference metrics.cpp:292)"); return -1; }
  double *xx[3][3];
  for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) xx[a][c] = calloc(n, sizeof(double));
  for (int p = 0; p < 9 * n; p++) b->met[p] = 0.0;
  for (int ph = 0; ph < o->ndim; ph++)
    for (int cp = 0; cp < o->ndim; cp++) {
      if (deriv_xi(o, b, b->xyz[ph], cp)) return -1;
      memcpy(xx[ph][cp], o->dflux, n * sizeof(double));
    }
  double *m = b->met;
  for (int p = 0; p < n; p++) {
    m[0 * n + p] = xx[1][1][p];   /* xi_x  =  y_eta */
    m[1 * n + p] = -xx[0][1][p];  /* xi_y  = -x_eta */
    m[3 * n + p] = -xx[1][0][p];  /* eta_x = -y_xi */
    m[4 * n + p] = xx[0][0][p];   /* eta_y =  x_xi */
  }
  for (int p = 0; p < n; p++) {
    double inv = xx[0][0][p] * xx[1][1][p] - xx[0][1][p] * xx[1][0][p];
    b->jac[p] = 1.0 / inv;
  }
  for (int p = 0; p < 9 * n; p++) b->minv[p] = 0.0;
  for (int p = 0; p < n; p++) {
    b->minv[0 * n + p] = m[4 * n + p];  /* [X][XI]  = eta_y */
    b->minv[1 * n + p] = -m[1 * n + p]; /* [X][ETA] = -xi_y */
    b->minv[3 * n + p] = -m[3 * n + p]; /* [Y][XI]  = -eta_x */
    b->minv[4 * n + p] = m[0 * n + p];  /* [Y][ETA] = xi_x */
  }
  for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) free(xx[a][c]);
  return 0;
}

/* State::compute_auxiliary_variables_linear_Euler, state.cpp:701-736 */
static void compute_aux(orc_t *o, blk_t *b, double **y) {
  double g = o->gamma, gi = 1.0 / g, gg = g / (g - 1.0);
  for (int p = 0; p < b->n; p++) {
    double sbar = b->mean[0][p], pbar = b->mean[4][p];
    double pbi = 1.0 / pbar;
    double rhobar = pow(g * pbar, gi) * exp(-sbar);
    double rbi = 1.0 / rhobar;
    double Tbar = pbar * gg * rbi;
    b->aux[1][p] = rhobar;
    b->aux[0][p] = y[IP][p] * gi * pbi - y[IS_][p];
    b->aux[0][p] *= rhobar;
    b->aux[3][p] = Tbar;
    b->aux[2][p] = y[IP][p] * pbi - b->aux[0][p] * rbi;
    b->aux[2][p] *= Tbar;
  }
}

/* additionalInit_bufferZones, geometry.cpp:1028-1142: sigma = C (d_in /
 * (d_in + d_out))^n with brute-force minimum distances to the zone's inflow
 * and outflow faces (IN is the face away from the block boundary). */
static void compute_sigma(orc_t *o, blk_t *b, bnd_t *z) {
  int nz[3], cnt = 1;
  for (int d = 0; d < 3; d++) { nz[d] = z->ie[d] - z->is[d] + 1; cnt *= nz[d]; }
  z->sigma = calloc(cnt, sizeof(double));
  /* face index along dir: IN = far side from the boundary, OUT = boundary side */
  int fin = z->end == 0 ? z->ie[z->dir] : z->is[z->dir];
  int fout = z->end == 0 ? z->is[z->dir] : z->ie[z->dir];
  int o1 = z->dir == 0 ? 1 : 0, o2 = z->dir == 2 ? 1 : 2;
  for (int k = 0; k < nz[2]; k++)
    for (int j = 0; j < nz[1]; j++)
      for (int i = 0; i < nz[0]; i++) {
        int gi = i + z->is[0], gj = j + z->is[1], gk = k + z->is[2];
        int l0 = idx(b, gi, gj, gk);
        double d2[2];
        for (int side = 0; side < 2; side++) {
          int f = side == 0 ? fin : fout;
          double dmin = fabs(-1e9);
          for (int c2 = z->is[o2]; c2 <= z->ie[o2]; c2++)
            for (int c1 = z->is[o1]; c1 <= z->ie[o1]; c1++) {
              int q[3];
              q[z->dir] = f; q[o1] = c1; q[o2] = c2;
              int lb = idx(b, q[0], q[1], q[2]);
              double dist = 0.0;
              for (int d = 0; d < 3; d++) dist += pow(b->xyz[d][l0] - b->xyz[d][lb], 2);
              dmin = dmin < dist ? dmin : dist;
            }
          d2[side] = sqrt(dmin);
        }
        double rel = d2[0] / (d2[0] + d2[1]);
        z->sigma[(k * nz[1] + j) * nz[0] + i] = o->sponge_constant * pow(rel, (double)o->sponge_order);
      }
}

/* ---------------------------------------------------------------- physics */

/* compute_RHS_linearized_Euler, spatial:2053-2130 */
static int rhs_lee(orc_t *o, blk_t *b, double **y, double **rhs) {
  int n = b->n, nd = o->ndim;
  double *df = o->dflux;
  double g = o->gamma;
  for (int v = 0; v < o->nv; v++) for (int p = 0; p < n; p++) rhs[v][p] = 0.0;
  for (int d = 0; d < nd; d++) {
    if (deriv_xyz(o, b, y[IS_], d)) return -1;
    for (int p = 0; p < n; p++) rhs[IS_][p] -= b->mean[IUX + d][p] * df[p];
    for (int p = 0; p < n; p++) rhs[IS_][p] -= y[IUX + d][p] * b->mgrad[IS_ * nd + d][p];
  }
  for (int vel = 0; vel < nd; vel++) {
    for (int d = 0; d < nd; d++) {
      if (deriv_xyz(o, b, y[IUX + vel], d)) return -1;
      for (int p = 0; p < n; p++) rhs[IUX + vel][p] -= b->mean[IUX + d][p] * df[p];
      for (int p = 0; p < n; p++)
        rhs[IUX + vel][p] -= (y[IUX + d][p] + b->aux[0][p] / b->aux[1][p] * b->mean[IUX + d][p])
                             * b->mgrad[(IUX + vel) * nd + d][p];
    }
    if (deriv_xyz(o, b, y[IP], vel)) return -1;
    for (int p = 0; p < n; p++) rhs[IUX + vel][p] -= df[p] / b->aux[1][p];
  }
  for (int d = 0; d < nd; d++) {
    if (deriv_xyz(o, b, y[IP], d)) return -1;
    for (int p = 0; p < n; p++) rhs[IP][p] -= b->mean[IUX + d][p] * df[p];
    for (int p = 0; p < n; p++) rhs[IP][p] -= y[IUX + d][p] * b->mgrad[IP * nd + d][p];
    if (deriv_xyz(o, b, y[IUX + d], d)) return -1;
    for (int p = 0; p < n; p++) rhs[IP][p] -= g * b->mean[IP][p] * df[p];
    for (int p = 0; p < n; p++) rhs[IP][p] -= g * y[IP][p] * b->mgrad[(IUX + d) * nd + d][p];
  }
  /* additional terms due to axisymmetry (spatial:2119-2126): IVAR_UR = 2,
   * RDIR = y, DUMMY_SMALL = 1e-99 (param.h:40-74) */
  if (o->axisym)
    for (int p = 0; p < n; p++)
      rhs[IP][p] -= g * (b->mean[IP][p] * y[IUY][p] + y[IP][p] * b->mean[IUY][p]) / (b->xyz[1][p] + 1e-99);
  /* passive scalars, compute_RHS_linearized_Euler_scalar (spatial:2134-2158),
   * called after the LEE terms (spatial:1977-1979):
   * rhs_Z -= ubar_i dZ'/dx_i, then rhs_Z -= u'_i dZbar/dx_i, direction by direction */
  for (int v = NV; v < o->nv; v++)
    for (int d = 0; d < nd; d++) {
      if (deriv_xyz(o, b, y[v], d)) return -1;
      for (int p = 0; p < n; p++) rhs[v][p] -= b->mean[IUX + d][p] * df[p];
      for (int p = 0; p < n; p++) rhs[v][p] -= y[IUX + d][p] * b->mgrad[v * nd + d][p];
    }
  return 0;
}

/* apply_physicalSource, source.cpp:22-107 (+ the monopole extension) */
static void physical_source(orc_t *o, blk_t *b, double **rhs, double time) {
  int n = b->n;
  if (o->src_kind == SRC_TWOCYL) {
    double mf = o->gamma - 1.0;
    double ef = -log(2.0) / (0.2 * 0.2);
    double sin_wt = sin(4.0 * o->twopi * time);
    double x0 = 0.0, y0 = 0.0;
    double ramp = 1.0, ramp_time = 16.0 / 4.0;
    if (time < ramp_time) ramp = pow(sin(o->pi / 2.0 * time / ramp_time), 2);
    for (int p = 0; p < n; p++) {
      double xs = b->xyz[0][p] - x0, ys = b->xyz[1][p] - y0;
      rhs[IP][p] += ramp * mf * exp(ef * (xs * xs + ys * ys)) * sin_wt;
    }
  } else if (o->src_kind == SRC_JET) {
    double amplitude = 0.001, frequency = 76.0;
    double Bx = 0.04 * log(2.0), By = 0.32 * log(2.0);
    double bref = 1.3, R = 287.0, g = o->gamma, rhoinf = 1.184, Tinf = 300.0;
    double cinf = sqrt(g * R * Tinf);
    amplitude /= rhoinf * pow(cinf, 2) / (bref / cinf);
    frequency *= o->inv_twopi * (bref / cinf);
    Bx *= pow(bref, 2);
    By *= pow(bref, 2);
    double w = frequency * o->twopi;
    double acos_wt = amplitude * cos(w * time);
    for (int p = 0; p < n; p++) {
      double xl = b->xyz[0][p], yl = b->xyz[1][p];
      rhs[IP][p] += exp(-(Bx * xl * xl + By * yl * yl)) * acos_wt;
    }
  } else if (o->src_kind == SRC_MONOPOLE) {
    /* src = {amplitude, omega, x0, y0, z0, halfwidth}; evaluated at the same
     * (t_{n+1} + c_i dt) times as the reference's sources */
    double A = o->src[0], om = o->src[1], hw = o->src[5];
    double k = (g_plant == 9 ? -1.0 : -log(2.0)) / (hw * hw);
    double a_sin = A * (g_plant == 8 ? cos(om * time) : sin(om * time));
    for (int p = 0; p < n; p++) {
      double dx = b->xyz[0][p] - o->src[2], dy = b->xyz[1][p] - o->src[3], dz = b->xyz[2][p] - o->src[4];
      rhs[IP][p] += exp(k * (dx * dx + dy * dy + dz * dz)) * a_sin;
    }
  }
}

/* bc_dirichlet_harmonicwave, boundary_condition.cpp:232-344, for LEE and
 * the one wave type that passes the reference's deck check, WAVE_ENTROPY
 * (HOMOGENEOUS or GAUSSIAN_ROUNDJET): data = {s', 0, 0, 0, 0} at a point
 * that is not a hole, all 0 at a hole. */
static void harmonic_data(const orc_t *o, const blk_t *b, int l0, double time, double *data) {
  static const int dir_other[3][2] = {{1, 2}, {0, 2}, {0, 1}}; /* param.h:133 */
  double rho_0 = 1.0, c_0 = 1.0;
  double wavenumber = o->twopi / o->hw_wavelength;
  for (int v = 0; v < o->nv; v++) data[v] = 0.0;
  if (b->iblank[l0] == BLANKED) return;
  double t1 = b->xyz[dir_other[o->hw_dir][0]][l0], t2 = b->xyz[dir_other[o->hw_dir][1]][l0];
  double pressure = 0.0, velocity = 0.0;
  double entropy = o->hw_amp * (g_plant == 10 ? 1.0 + cos(wavenumber * c_0 * time)
                                              : 1.0 + sin(wavenumber * c_0 * time));
  if (o->hw_shape == WAVEFORM_GAUSSIAN_ROUNDJET) {
    double fac_Gaussian = -log(2.0) / pow(o->hw_halfwidth, 2);
    entropy *= exp(fac_Gaussian * (pow(t1, 2) + pow(t2, 2)));
  }
  data[IS_] = entropy;
  data[IUX + o->hw_dir] = velocity / (rho_0 * c_0);
  data[IP] = pressure;
}

/* apply_bufferZone_Freund_ambient, source.cpp:174-201 (sol_ref = 0), and
 * apply_bufferZone_Freund_harmonicWave, source.cpp:264-320 */
static void sponges(orc_t *o, blk_t *b, double **y, double **rhs, double time) {
  for (int z = 0; z < b->nbnd; z++) {
    bnd_t *zb = &b->bnd[z];
    if (zb->type != SPONGE_AMBIENT && zb->type != SPONGE_HARMONIC) continue;
    int nz0 = zb->ie[0] - zb->is[0] + 1, nz1 = zb->ie[1] - zb->is[1] + 1;
    for (int k = zb->is[2]; k <= zb->ie[2]; k++)
      for (int j = zb->is[1]; j <= zb->ie[1]; j++)
        for (int i = zb->is[0]; i <= zb->ie[0]; i++) {
          int lb = ((k - zb->is[2]) * nz1 + (j - zb->is[1])) * nz0 + (i - zb->is[0]);
          int l0 = idx(b, i, j, k);
          double s = zb->sigma[lb];
          double ref[MAXV] = {0.0};
          if (zb->type == SPONGE_HARMONIC) harmonic_data(o, b, l0, time, ref);
          for (int v = 0; v < o->nv; v++) {
            double relax = -s * (y[v][l0] - ref[v]);
            rhs[v][l0] += relax;
          }
        }
  }
}

/* spatial::compute_RHS, spatial:1963-1990 */
static int compute_rhs_all(orc_t *o, double ***y, double ***rhs, double time) {
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    if (rhs_lee(o, b, y[ib], rhs[ib])) return -1;
    physical_source(o, b, rhs[ib], time);
    sponges(o, b, y[ib], rhs[ib], time);
  }
  return 0;
}

/* bc::apply_BC, boundary_condition.cpp:46-209 with bc_neumann :348-444,
 * bc_wall_slip_kinematic :448-494, bc_dirichlet_allzero, bc_centerline */
static void apply_bc(orc_t *o, blk_t *b, double **y, double time) {
  for (int ib = 0; ib < b->nbnd; ib++) {
    bnd_t *bc = &b->bnd[ib];
    if (bc->type >= SPONGE_AMBIENT) continue;
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) { lo[d] = bc->is[d]; hi[d] = bc->ie[d]; }
    lo[bc->dir] = hi[bc->dir] = bc->end == 0 ? bc->is[bc->dir] : bc->ie[bc->dir];
    int mf = stride_of(b, bc->dir);
    int vars[MAXV], nvu = 0;
    int neumann = bc->type == BC_NEUMANN || bc->type == BC_WALL_SLIP || (bc->type >= BC_CL_X && bc->type <= 903);
    /* wall-slip: the non-velocity variables (s', p' and the scalars) take
     * the Neumann data (boundary_condition.cpp:124-136) */
    if (bc->type == BC_WALL_SLIP) {
      vars[nvu++] = IS_; vars[nvu++] = IP;
      if (g_plant != 12)
        for (int v = NV; v < o->nv; v++) vars[nvu++] = v;
    } else for (int v = 0; v < o->nv; v++) vars[nvu++] = v;
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) {
          int l0 = idx(b, i, j, k);
          if (b->iblank[l0] == BLANKED) continue;
          double data[MAXV];
          for (int v = 0; v < o->nv; v++) data[v] = 1e9 * (double)(bc->dir + 1); /* -DUMMY_DOUBLE*(dir+1) */
          if (bc->type == BC_DIRICHLET_ALLZERO) {
            for (int v = 0; v < o->nv; v++) data[v] = 0.0;
          }
          if (bc->type == BC_DIRICHLET_HARMONIC) harmonic_data(o, b, l0, time, data);
          if (neumann) {
            const op_t *op = &o->fd;
            for (int q = 0; q < nvu; q++) {
              int v = vars[q];
              data[v] = 0.0;
              if (bc->end == 0) {
                for (int s = 1; s < op->sb; s++) data[v] -= op->bnd[0][s] * y[v][l0 + mf * s];
                data[v] /= op->bnd[0][0];
              } else {
                for (int s = 1; s < op->sb; s++) data[v] += op->bnd[0][s] * y[v][l0 - mf * s];
                data[v] /= -op->bnd[0][0];
              }
            }
          }
          if (bc->type == BC_WALL_SLIP) {
            double vc[3], vk[3];
            for (int d = 0; d < 3; d++) vc[d] = y[IUX + d][l0];
            for (int c = 0; c < o->ndim; c++) {
              vk[c] = 0.0;
              for (int d = 0; d < o->ndim; d++) vk[c] += b->met[(3 * c + d) * b->n + l0] * vc[d];
            }
            vk[bc->dir] = 0.0;
            for (int d = 0; d < o->ndim; d++) {
              vc[d] = 0.0;
              for (int c = 0; c < o->ndim; c++) vc[d] += b->minv[(3 * d + c) * b->n + l0] * vk[c];
              vc[d] *= b->jac[l0];
            }
            for (int d = 0; d < 3; d++) data[IUX + d] = vc[d];
          }
          if (bc->type >= BC_CL_X && bc->type <= BC_CL_Z) data[IUX + (bc->type - BC_CL_X)] = 0.0;
          /* BC_CENTERLINE_AXISYM: Neumann, then u_r = 0 (bc:150-156, 534-541) */
          if (bc->type == 903) data[IUY] = 0.0;
          for (int v = 0; v < o->nv; v++) y[v][l0] = data[v];
        }
  }
}

/* overset::interpolate (OVERTURE), overset.cpp:820-845, 886-974, 1241-1278:
 * every weighted sum is formed before any receiver is overwritten. */
static void overset_interp(orc_t *o, double ***y) {
  ovs_t *ov = &o->ov;
  for (int r = 0; r < ov->nrecv; r++) {
    blk_t *db = &o->b[ov->db[r]];
    const int *w = ov->w[ov->db[r]];
    const double *c = ov->coef + (size_t)r * ov->maxst;
    double acc[MAXV] = {0};
    int s = 0;
    for (int w3 = 0; w3 < w[2]; w3++)
      for (int w2 = 0; w2 < w[1]; w2++)
        for (int w1 = 0; w1 < w[0]; w1++) {
          int l1 = ov->dl[r] + (w1 + w2 * db->ni + w3 * db->ni * db->nj);
          for (int v = 0; v < o->nv; v++) acc[v] += c[s] * y[ov->db[r]][v][l1];
          s++;
        }
    for (int v = 0; v < o->nv; v++) ov->buf[r * o->nv + v] = acc[v];
  }
  for (int r = 0; r < ov->nrecv; r++)
    for (int v = 0; v < o->nv; v++) y[ov->rb[r]][v][ov->rl[r]] = ov->buf[r * o->nv + v];
}

/* spatial::update_boundary, spatial:2162-2187 */
static void update_boundary(orc_t *o, double ***y, double time) {
  for (int ib = 0; ib < o->nblk; ib++) apply_bc(o, &o->b[ib], y[ib], time);
  if (o->do_overset) overset_interp(o, y);
}

/* spatial::apply_filter, spatial:2191-2246 */
static int apply_filter_block(orc_t *o, blk_t *b, double **y) {
  int n = b->n;
  for (int v = 0; v < (g_plant == 11 ? NV : o->nv); v++) {
    for (int p = 0; p < n; p++) o->func0[p] = y[v][p];
    for (int p = 0; p < n; p++) o->func[p] = o->func0[p];
    for (int dd = 0; dd < o->ndim; dd++) {
      int d = g_plant == 6 ? o->ndim - 1 - dd : dd;
      for (int p = 0; p < n; p++) o->dflux[p] = o->func[p];
      if (apply_op(b, &o->flt, 1, o->func, o->dflux, d)) return -1;
      for (int p = 0; p < n; p++) o->func[p] = o->dflux[p];
    }
    double rem = 1.0 - o->filter_blend;
    for (int p = 0; p < n; p++) y[v][p] = rem * o->func0[p] + o->filter_blend * o->func[p];
  }
  return 0;
}

/* RungeKutta::integrate, temporal_discretization.cpp:177-252 */
static int rk_integrate(orc_t *o, double tn, double h) {
  double **ynext[MAXB], **yint[MAXB], **rhs[MAXB];
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    ynext[ib] = b->ynp1; yint[ib] = b->interim; rhs[ib] = b->rhs;
    for (int v = 0; v < o->nv; v++)
      for (int p = 0; p < b->n; p++) { b->ynp1[v][p] = b->yn[v][p]; b->interim[v][p] = b->yn[v][p]; }
  }
  int last = o->rk_stages - 1;
  for (int st = 0; st <= last; st++) {
    if (compute_rhs_all(o, yint, rhs, tn + o->c_i[st] * h)) return -1;
    for (int ib = 0; ib < o->nblk; ib++) {
      blk_t *b = &o->b[ib];
      for (int p = 0; p < b->n; p++) {
        if (b->iblank[p] != BLANKED) {
          for (int v = 0; v < o->nv; v++) {
            b->rhs[v][p] *= h;
            b->ynp1[v][p] += o->b_i[st] * b->rhs[v][p];
            if (st < last) b->interim[v][p] = b->yn[v][p] + o->a_ij[st + 1][st] * b->rhs[v][p];
          }
        } else {
          for (int v = 0; v < o->nv; v++) { b->rhs[v][p] = 0.0; b->ynp1[v][p] = 0.0; b->interim[v][p] = 0.0; }
        }
      }
    }
    int nx = st + 1 < last ? st + 1 : last;
    double tb = tn + o->c_i[nx] * h; /* boundary time tn + c_{min(i+1,last)} h (temporal:236-246) */
    if (st < last) {
      update_boundary(o, yint, tb);
      for (int ib = 0; ib < o->nblk; ib++) compute_aux(o, &o->b[ib], o->b[ib].interim);
    } else {
      update_boundary(o, ynext, tb);
      for (int ib = 0; ib < o->nblk; ib++) compute_aux(o, &o->b[ib], o->b[ib].ynp1);
    }
  }
  return 0;
}

/* temporal::time_advance, temporal:334-376 */
static int time_advance(orc_t *o) {
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    for (int v = 0; v < o->nv; v++)
      for (int p = 0; p < b->n; p++) { b->yn[v][p] = b->sol_old[v][p]; b->ynp1[v][p] = -1e9; }
  }
  if (rk_integrate(o, g_plant == 2 ? o->time_sol - o->dt : o->time_sol, o->dt)) return -1;
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    if (o->do_filter && apply_filter_block(o, b, b->ynp1)) return -1;
    for (int v = 0; v < o->nv; v++)
      for (int p = 0; p < b->n; p++) if (b->iblank[p] == BLANKED) b->ynp1[v][p] = 0.0;
    for (int v = 0; v < o->nv; v++) memcpy(b->sol[v], b->ynp1[v], b->n * sizeof(double));
    if (g_plant == 3) compute_aux(o, b, b->sol);
  }
  return 0;
}

/* temporal::compute_CFL, temporal:397-450 (the convective term goes through
 * C's int abs(int): each product is truncated toward zero, :425) */
static double compute_cfl(orc_t *o) {
  double cmax = -1e99;
  double gm1 = o->gamma - 1.0;
  int nd = o->ndim;
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    int n = b->n;
    for (int p = 0; p < n; p++) {
      if (b->iblank[p] != NOTBLANKED) continue;
      double vc[3], vk[3];
      for (int d = 0; d < 3; d++) vc[d] = b->mean[IUX + d][p];
      for (int d = nd; d < 3; d++) vc[d] = 0.0;
      for (int c = 0; c < nd; c++) {
        vk[c] = 0.0;
        for (int d = 0; d < nd; d++) vk[c] += b->met[(3 * c + d) * n + p] * vc[d];
      }
      double cfl = 0.0;
      for (int c = 0; c < nd; c++)
        for (int d = 0; d < nd; d++) {
          if (g_plant == 1) cfl += fabs(vk[c] * b->met[(3 * c + d) * n + p]);
          else cfl += abs((int)(vk[c] * b->met[(3 * c + d) * n + p]));
        }
      double fac = 0.0;
      for (int c = 0; c < nd; c++)
        for (int d = 0; d < nd; d++) fac += b->met[(3 * c + d) * n + p] * b->met[(3 * c + d) * n + p];
      double sos = sqrt(gm1 * b->aux[3][p]);
      cfl += sos * sqrt(fac);
      cfl *= o->dt * b->jac[p];
      cmax = cmax > cfl ? cmax : cfl;
    }
  }
  return cmax;
}

/* ------------------------------------------------------------- public API */

void *orc_create(int nblk, int ndim, double gamma, int fd_order, int filter_order, double filter_strength,
                 double filter_blend, int rk_stages, int sponge_order, double sponge_constant) {
  if (nblk < 1 || nblk > MAXB) { set_err("bad block count"); return NULL; }
  orc_t *o = calloc(1, sizeof(orc_t));
  o->nblk = nblk;
  o->nv = NV;
  o->ndim = ndim;
  o->gamma = gamma;
  o->pi = acos(-1.0);
  o->twopi = 2.0 * o->pi;
  o->inv_twopi = 1.0 / o->twopi;
  if (init_fd(&o->fd, fd_order)) { free(o); return NULL; }
  o->do_filter = filter_order > 0;
  if (o->do_filter) {
    /* the standard central filter always runs at full strength (input.cpp:50-53);
     * SFO11P takes the deck's STRENGTH */
    if (init_filter(&o->flt, filter_order, filter_order == 11 ? filter_strength : 1.0)) { free(o); return NULL; }
  } else {
    o->flt.ghost = 0;
  }
  o->filter_blend = filter_blend;
  o->rk_stages = rk_stages;
  if (rk_stages == 4) {
    o->b_i[0] = 1.0 / 6.0; o->b_i[1] = 1.0 / 3.0; o->b_i[2] = 1.0 / 3.0; o->b_i[3] = 1.0 / 6.0;
    o->c_i[0] = 0.0; o->c_i[1] = 1.0 / 2.0; o->c_i[2] = 1.0 / 2.0; o->c_i[3] = 1.0;
    o->a_ij[1][0] = 1.0 / 2.0; o->a_ij[2][1] = 1.0 / 2.0; o->a_ij[3][2] = 1.0;
  } else if (rk_stages == 2) {
    o->b_i[0] = 1.0 / 2.0; o->b_i[1] = 1.0 / 2.0;
    o->c_i[0] = 0.0; o->c_i[1] = 1.0;
    o->a_ij[1][0] = 1.0;
  } else { set_err("RK must have 2 or 4 stages"); free(o); return NULL; }
  o->sponge_order = sponge_order;
  o->sponge_constant = sponge_constant;
  o->dt = 1e-6;
  return o;
}

static double *dalloc(int n) { return calloc(n, sizeof(double)); }

int orc_set_block(void *h, int ib, int ni, int nj, int nk, const double *xyz, const double *mean) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  b->ni = ni; b->nj = nj; b->nk = nk;
  b->nd[0] = ni; b->nd[1] = nj; b->nd[2] = nk;
  int n = b->n = ni * nj * nk;
  for (int d = 0; d < 3; d++) { b->xyz[d] = dalloc(n); memcpy(b->xyz[d], xyz + (size_t)d * n, n * sizeof(double)); }
  b->met = dalloc(9 * n);
  b->minv = dalloc(9 * n);
  b->jac = dalloc(n);
  b->iblank = calloc(n, sizeof(int));
  for (int v = 0; v < o->nv; v++) {
    b->mean[v] = dalloc(n);
    memcpy(b->mean[v], mean + (size_t)v * n, n * sizeof(double));
    b->sol[v] = dalloc(n); b->sol_old[v] = dalloc(n);
    b->rhs[v] = dalloc(n); b->interim[v] = dalloc(n); b->yn[v] = dalloc(n); b->ynp1[v] = dalloc(n);
  }
  for (int v = 0; v < o->nv * 3; v++) b->mgrad[v] = dalloc(n);
  for (int v = 0; v < 4; v++) b->aux[v] = dalloc(n);
  if (n > o->nmax) {
    o->nmax = n;
    free(o->dflux); free(o->dflux_sum); free(o->flux); free(o->func0); free(o->func);
    o->dflux = dalloc(n); o->dflux_sum = dalloc(n); o->flux = dalloc(n); o->func0 = dalloc(n); o->func = dalloc(n);
  }
  set_extents(o, b);
  return 0;
}

/* PHYSICAL_MODEL = LEE_SCALAR with NUMBER_SCALAR = ns (input.cpp:113-119,
 * 768-774): ns passive scalars Z'_k after p' (state.cpp:336-372); call
 * before orc_set_block. */
int orc_set_scalars(void *h, int ns) {
  orc_t *o = h;
  if (ns < 0 || NV + ns > MAXV) { set_err("NUMBER_SCALAR must be 0 to 4."); return -1; }
  o->nv = NV + ns;
  return 0;
}

/* AXISYMMETRY = TRUE with AXIS_OF_SYMMETRY = X (input.cpp:105-108, 523-531) */
int orc_set_axisym(void *h, int on) {
  orc_t *o = h;
  if (on && o->ndim != 2) { set_err("Axisymmetry requires 2-D."); return -1; }
  o->axisym = on;
  return 0;
}

/* A BC-file region (io.cpp:1158-1268): type, dir, end (0 '+' / 1 '-'),
 * block-index box is[3]..ie[3] after resolving negative indices. */
int orc_add_boundary(void *h, int ib, int type, int dir, int end, const int *is, const int *ie) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  if (b->nbnd >= 32) { set_err("too many boundaries"); return -1; }
  bnd_t *z = &b->bnd[b->nbnd++];
  memset(z, 0, sizeof *z);
  z->type = type; z->dir = dir; z->end = end;
  for (int d = 0; d < 3; d++) { z->is[d] = is[d]; z->ie[d] = ie[d]; }
  return 0;
}

void orc_set_source(void *h, int kind, const double *params) {
  orc_t *o = h;
  o->src_kind = kind;
  if (params) memcpy(o->src, params, sizeof o->src);
}

/* HARMONIC_WAVE type, SHAPE, DIRECTION (0 x, 1 y, 2 z), AMPLITUDE,
 * WAVELENGTH, HALFWIDTH */
void orc_set_harmonic(void *h, int wave, int shape, int dir, double amp, double wavelength, double halfwidth) {
  orc_t *o = h;
  o->hw_wave = wave; o->hw_shape = shape; o->hw_dir = dir;
  o->hw_amp = amp; o->hw_wavelength = wavelength; o->hw_halfwidth = halfwidth;
}

void orc_set_timestep(void *h, int fix_cfl, double value) {
  orc_t *o = h;
  o->fix_cfl = fix_cfl;
  if (fix_cfl) { o->cfl_target = value; o->dt = 1e-6; }
  else { o->dt = value; o->cfl_target = 0.0; }
}

/* Overset tables: receivers (block, 1-D index), donors (block, lower-left
 * 1-D index), raw coefficients (nrecv x stencil, w1 fastest), widths per
 * block, holes (block, 1-D index).  Coefficients are normalised like
 * initialize_donor (overset.cpp:516-526): c *= 1 / sum(c). */
int orc_set_overset(void *h, int nrecv, const int *rb, const int *rl, const int *db, const int *dl,
                    const int *widths /* nblk*3 */, const double *coef, int maxst, int nhole,
                    const int *hb, const int *hl) {
  orc_t *o = h;
  ovs_t *ov = &o->ov;
  o->do_overset = 1;
  ov->nrecv = nrecv;
  ov->maxst = maxst;
  ov->rb = malloc(nrecv * sizeof(int)); ov->rl = malloc(nrecv * sizeof(int));
  ov->db = malloc(nrecv * sizeof(int)); ov->dl = malloc(nrecv * sizeof(int));
  ov->coef = malloc((size_t)nrecv * maxst * sizeof(double));
  ov->buf = malloc((size_t)nrecv * o->nv * sizeof(double));
  memcpy(ov->rb, rb, nrecv * sizeof(int)); memcpy(ov->rl, rl, nrecv * sizeof(int));
  memcpy(ov->db, db, nrecv * sizeof(int)); memcpy(ov->dl, dl, nrecv * sizeof(int));
  for (int ib = 0; ib < o->nblk; ib++) for (int d = 0; d < 3; d++) ov->w[ib][d] = widths[ib * 3 + d];
  for (int r = 0; r < nrecv; r++) {
    const int *w = ov->w[db[r]];
    int ns = w[0] * w[1] * w[2];
    const double *c = coef + (size_t)r * maxst;
    double *cn = ov->coef + (size_t)r * maxst;
    double sum = 0.0;
    for (int s = 0; s < ns; s++) { cn[s] = c[s]; sum += cn[s]; }
    sum = 1.0 / sum;
    for (int s = 0; s < ns; s++) cn[s] *= sum;
  }
  /* override_iblank_4holeCells then _4interpolatedCells (overset.cpp:683-790) */
  for (int q = 0; q < nhole; q++) o->b[hb[q]].iblank[hl[q]] = BLANKED;
  for (int r = 0; r < nrecv; r++) o->b[rb[r]].iblank[rl[r]] = INTERPOLATED;
  return 0;
}

/* The part of simulation::initialize after the solution and base state are
 * read (simulation.cpp:164-240): aux, sponge strengths, metrics (before any
 * hole exists), then -- with overset already marked -- mean gradients. */
int orc_init(void *h, const double *sol0_all, int step0, double time0) {
  orc_t *o = h;
  size_t off = 0;
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    for (int v = 0; v < o->nv; v++) memcpy(b->sol[v], sol0_all + off + (size_t)v * b->n, b->n * sizeof(double));
    off += (size_t)o->nv * b->n;
    compute_aux(o, b, b->sol);
    for (int z = 0; z < b->nbnd; z++)
      if (b->bnd[z].type == SPONGE_AMBIENT || b->bnd[z].type == SPONGE_HARMONIC) compute_sigma(o, b, &b->bnd[z]);
    /* metrics are computed with all points fluid: stash iblank */
    int *ibk = b->iblank;
    int *zero = calloc(b->n, sizeof(int));
    b->iblank = zero;
    if (compute_metrics(o, b)) { b->iblank = ibk; free(zero); return -1; }
    b->iblank = ibk;
    free(zero);
  }
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    for (int v = 0; v < o->nv; v++)
      for (int d = 0; d < o->ndim; d++) {
        if (deriv_xyz(o, b, b->mean[v], d)) return -1;
        memcpy(b->mgrad[v * o->ndim + d], o->dflux, b->n * sizeof(double));
      }
  }
  if (g_plant == 5)
    for (int ib = 0; ib < o->nblk; ib++)
      for (int v = 0; v < o->nv * o->ndim; v++) memset(o->b[ib].mgrad[v], 0, o->b[ib].n * sizeof(double));
  o->time_step = step0;
  o->time_sol = time0;
  return 0;
}

/* One iteration of solver::solve, solver.cpp:41-89 */
int orc_step(void *h) {
  orc_t *o = h;
  for (int ib = 0; ib < o->nblk; ib++)
    for (int v = 0; v < o->nv; v++) memcpy(o->b[ib].sol_old[v], o->b[ib].sol[v], o->b[ib].n * sizeof(double));
  o->time_step++;
  double cmax = compute_cfl(o);
  double dt = o->dt;
  if (o->fix_cfl) {
    double ratio = o->cfl_target / cmax;
    dt *= ratio;
    o->dt = dt;
  }
  o->time_sol += dt;
  return time_advance(o);
}

int orc_run(void *h, int nsteps) {
  for (int s = 0; s < nsteps; s++) if (orc_step(h)) return -1;
  return 0;
}

double orc_time(void *h) { return ((orc_t *)h)->time_sol; }
double orc_dt(void *h) { return ((orc_t *)h)->dt; }
int orc_time_step(void *h) { return ((orc_t *)h)->time_step; }
double orc_cfl(void *h) { return compute_cfl((orc_t *)h); }

void orc_get_solution(void *h, int ib, double *out) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  for (int v = 0; v < o->nv; v++) memcpy(out + (size_t)v * b->n, b->sol[v], b->n * sizeof(double));
}

void orc_set_solution(void *h, int ib, const double *in) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  for (int v = 0; v < o->nv; v++) memcpy(b->sol[v], in + (size_t)v * b->n, b->n * sizeof(double));
  compute_aux(o, b, b->sol);
}

void orc_set_time(void *h, int step, double time, double dt) {
  orc_t *o = h;
  o->time_step = step;
  o->time_sol = time;
  o->dt = dt;
}

/* metrics[l][i] (9n), metricsInverse (9n), J (n) */
void orc_get_metrics(void *h, int ib, double *met, double *minv, double *jac) {
  blk_t *b = &((orc_t *)h)->b[ib];
  if (met) memcpy(met, b->met, 9 * (size_t)b->n * sizeof(double));
  if (minv) memcpy(minv, b->minv, 9 * (size_t)b->n * sizeof(double));
  if (jac) memcpy(jac, b->jac, (size_t)b->n * sizeof(double));
}

void orc_get_mean_gradient(void *h, int ib, double *g) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  for (int v = 0; v < o->nv * o->ndim; v++) memcpy(g + (size_t)v * b->n, b->mgrad[v], b->n * sizeof(double));
}

void orc_get_aux(void *h, int ib, double *aux) {
  blk_t *b = &((orc_t *)h)->b[ib];
  for (int v = 0; v < 4; v++) memcpy(aux + (size_t)v * b->n, b->aux[v], b->n * sizeof(double));
}

void orc_get_iblank(void *h, int ib, int *ibk) {
  blk_t *b = &((orc_t *)h)->b[ib];
  memcpy(ibk, b->iblank, b->n * sizeof(int));
}

/* Sum of sponge strengths per point (for checks; the solver applies zones
 * one after the other). */
void orc_get_sigma(void *h, int ib, double *sig) {
  blk_t *b = &((orc_t *)h)->b[ib];
  for (int p = 0; p < b->n; p++) sig[p] = 0.0;
  for (int z = 0; z < b->nbnd; z++) {
    bnd_t *zb = &b->bnd[z];
    if ((zb->type != SPONGE_AMBIENT && zb->type != SPONGE_HARMONIC) || !zb->sigma) continue;
    int nz0 = zb->ie[0] - zb->is[0] + 1, nz1 = zb->ie[1] - zb->is[1] + 1;
    for (int k = zb->is[2]; k <= zb->ie[2]; k++)
      for (int j = zb->is[1]; j <= zb->ie[1]; j++)
        for (int i = zb->is[0]; i <= zb->ie[0]; i++)
          sig[idx(b, i, j, k)] += zb->sigma[((k - zb->is[2]) * nz1 + (j - zb->is[1])) * nz0 + (i - zb->is[0])];
  }
}

/* Unit entry points on one block's arrays (5n each). */
int orc_rhs(void *h, int ib, const double *y, double *rhs, double time, int with_aux_from_y) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  double *yy[MAXV], *rr[MAXV];
  for (int v = 0; v < o->nv; v++) { yy[v] = (double *)y + (size_t)v * b->n; rr[v] = rhs + (size_t)v * b->n; }
  if (with_aux_from_y) compute_aux(o, b, yy);
  if (rhs_lee(o, b, yy, rr)) return -1;
  physical_source(o, b, rr, time);
  sponges(o, b, yy, rr, time);
  return 0;
}

int orc_filter(void *h, int ib, double *y) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  double *yy[MAXV];
  for (int v = 0; v < o->nv; v++) yy[v] = y + (size_t)v * b->n;
  return apply_filter_block(o, b, yy);
}

/* One-directional operator on a single field: which = 0 derivative, 1 filter */
int orc_apply_dir(void *h, int ib, int which, const double *f, double *out, int dir) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  for (int p = 0; p < b->n; p++) out[p] = which ? f[p] : 0.0;
  return apply_op(b, which ? &o->flt : &o->fd, which, f, out, dir);
}

int orc_deriv_xyz(void *h, int ib, const double *f, double *out, int dir) {
  orc_t *o = h;
  blk_t *b = &o->b[ib];
  if (deriv_xyz(o, b, f, dir)) return -1;
  memcpy(out, o->dflux, b->n * sizeof(double));
  return 0;
}

void orc_set_iblank(void *h, int ib, const int *ibk) {
  blk_t *b = &((orc_t *)h)->b[ib];
  memcpy(b->iblank, ibk, b->n * sizeof(int));
}

void orc_destroy(void *h) {
  orc_t *o = h;
  if (!o) return;
  for (int ib = 0; ib < o->nblk; ib++) {
    blk_t *b = &o->b[ib];
    for (int d = 0; d < 3; d++) free(b->xyz[d]);
    free(b->met); free(b->minv); free(b->jac); free(b->iblank);
    for (int v = 0; v < o->nv; v++) {
      free(b->mean[v]); free(b->sol[v]); free(b->sol_old[v]); free(b->rhs[v]);
      free(b->interim[v]); free(b->yn[v]); free(b->ynp1[v]);
    }
    for (int v = 0; v < o->nv * 3; v++) free(b->mgrad[v]);
    for (int v = 0; v < 4; v++) free(b->aux[v]);
    for (int z = 0; z < b->nbnd; z++) free(b->bnd[z].sigma);
  }
  free(o->ov.rb); free(o->ov.rl); free(o->ov.db); free(o->ov.dl); free(o->ov.coef); free(o->ov.buf);
  free(o->dflux); free(o->dflux_sum); free(o->flux); free(o->func0); free(o->func);
  free(o);
}
