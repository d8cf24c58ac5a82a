// HIP kernels of the LEE time advance for gfx950 (CDNA4).
//
// Two families:
//  * plain_*  -- one kernel per reference operator (take_derivative_xyz, each
//    RHS term, sponge, source, RK update, filter sweep, blend, aux), FP64 with
//    floating-point contraction off and the reference's operation order, so
//    the GPU result equals the CPU reference bit for bit.  This is the
//    drop-in FP64 path and the yardstick for the fused path.
//  * fused_*  -- one LDS-tiled kernel per RK stage (both directional
//    derivatives of the four active variables, chain rule with FP64 metrics,
//    LEE right-hand side, source, sponge, RK update, aux density) and one
//    LDS-tiled kernel per step for the two filter sweeps + blend + hole
//    reset.  Templated on the bulk type (double or float).
//
// All global indices are bounds-checked against the block extents.
#include <hip/hip_runtime.h>

#include <cmath>

#include "lee_internal.h"
#include "lee_device.h"

namespace lee {

namespace {

// ------------------------------------------------------------------ plain

// One row of the reference operator at point p (coordinate c along a line of
// length len, memory stride st), every coefficient included, left to right.
__device__ double plain_row(const OpTable &op, const double *f, int p, int c, int len, int st, int code) {
#pragma clang fp contract(off)
  double acc = 0.0;
  if (code == CODE_INTERIOR) {
    for (int s = -op.half; s <= op.half; s++) {
      const int q = c + s;
      if (q >= 0 && q < len) acc += op.st[s + op.half] * f[p + s * st];
    }
  } else if (code == CODE_ZERO) {
    acc = 0.0;
  } else if (code < CODE_LEFT + op.nb) {
    const int k = code - CODE_LEFT;
    for (int s = 0; s < op.sb; s++) {
      const int q = c - k + s;
      if (q >= 0 && q < len) acc += op.bnd[k][s] * f[p + (s - k) * st];
    }
  } else {
    const int k = code - CODE_LEFT - op.nb;
    for (int s = 0; s < op.sb; s++) {
      const int q = c + k - s;
      if (q >= 0 && q < len) acc += op.sign_right * op.bnd[k][s] * f[p + (k - s) * st];
    }
  }
  return acc;
}

__global__ void __launch_bounds__(TPB) k_plain_op_dir(const double *__restrict__ f, double *__restrict__ df,
                                                      const int8_t *__restrict__ codes, int ni, int nj, int dir,
                                                      OpTable op) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  const int n = ni * nj;
  if (p >= n) return;
  const int i = p % ni, j = p / ni;
  const int c = dir == 0 ? i : j, len = dir == 0 ? ni : nj, st = dir == 0 ? 1 : ni;
  const int code = codes ? (int)codes[p] : default_code(c, len, op.nb);
  df[p] = plain_row(op, f, p, c, len, st, code);
}

// StandardCentral::take_derivative_xyz (spatial:1802-1826):
// out = (0 + m[xi][d] d/dxi + m[eta][d] d/deta) * J
__global__ void __launch_bounds__(TPB) k_plain_deriv_xyz(const double *__restrict__ f, double *__restrict__ out,
                                                         const double *__restrict__ met,
                                                         const double *__restrict__ jac,
                                                         const int8_t *__restrict__ codes, int ni, int nj, int dxyz,
                                                         OpTable op) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  const int n = ni * nj;
  if (p >= n) return;
  const int i = p % ni, j = p / ni;
  const int c0 = codes ? (int)codes[p] : default_code(i, ni, op.nb);
  const int c1 = codes ? (int)codes[n + p] : default_code(j, nj, op.nb);
  const double d0 = plain_row(op, f, p, i, ni, 1, c0);
  const double d1 = plain_row(op, f, p, j, nj, ni, c1);
  double sum = 0.0;
  sum += met[(size_t)(0 + dxyz) * n + p] * d0;
  sum += met[(size_t)(2 + dxyz) * n + p] * d1;
  out[p] = sum * jac[p];
}

// metrics::compute_ThomasLombard_2D + compute_Jacobian (metrics.cpp:107-129,
// 255-328).  xx = {x_xi, x_eta, y_xi, y_eta}.
__global__ void __launch_bounds__(TPB) k_metrics2d(const double *__restrict__ xx, double *__restrict__ met,
                                                   double *__restrict__ jac, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  const double x_xi = xx[p], x_eta = xx[(size_t)n + p], y_xi = xx[(size_t)2 * n + p], y_eta = xx[(size_t)3 * n + p];
  met[p] = y_eta;
  met[(size_t)n + p] = -x_eta;
  met[(size_t)2 * n + p] = -y_xi;
  met[(size_t)3 * n + p] = x_xi;
  const double inv = x_xi * y_eta - x_eta * y_xi;
  jac[p] = 1.0 / inv;
}

__global__ void __launch_bounds__(TPB) k_fill(double *__restrict__ a, double v, size_t n) {
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p < n) a[p] = v;
}

__global__ void __launch_bounds__(TPB) k_copy(const double *__restrict__ a, double *__restrict__ b, size_t n) {
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p < n) b[p] = a[p];
}

// rhs -= a * b
__global__ void __launch_bounds__(TPB) k_sub_ab(double *__restrict__ r, const double *__restrict__ a,
                                                const double *__restrict__ b, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p < n) r[p] -= a[p] * b[p];
}

// rhs -= (g * a) * b
__global__ void __launch_bounds__(TPB) k_sub_gab(double *__restrict__ r, double g, const double *__restrict__ a,
                                                 const double *__restrict__ b, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p < n) r[p] -= g * a[p] * b[p];
}

// rhs -= a / b
__global__ void __launch_bounds__(TPB) k_sub_div(double *__restrict__ r, const double *__restrict__ a,
                                                 const double *__restrict__ b, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p < n) r[p] -= a[p] / b[p];
}

// rhs -= (u' + rho'/rho_bar * u_bar) * grad   (spatial:2089-2092)
__global__ void __launch_bounds__(TPB) k_sub_velgrad(double *__restrict__ r, const double *__restrict__ up,
                                                     const double *__restrict__ rho,
                                                     const double *__restrict__ rhob,
                                                     const double *__restrict__ ub,
                                                     const double *__restrict__ grad, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p < n) r[p] -= (up[p] + rho[p] / rhob[p] * ub[p]) * grad[p];
}

// apply_physicalSource: rhs[P] += (s1 * g) * s2 over the source support
// compute_RHS_linearized_Euler's axisymmetric term (spatial:2119-2126):
// rhs_p -= gamma (pbar u_r' + p' ubar_r) / (r + 1e-99), with u_r = u_y and
// r = y (the axis of symmetry is x, input.cpp:523-531).
__global__ void __launch_bounds__(TPB) k_plain_axisym(double *__restrict__ rp, double gam,
                                                      const double *__restrict__ pb, const double *__restrict__ ur,
                                                      const double *__restrict__ pp, const double *__restrict__ urb,
                                                      const double *__restrict__ r, int n) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q >= n) return;
  const double DUMMY_SMALL = 1e-99;
  rp[q] -= gam * (pb[q] * ur[q] + pp[q] * urb[q]) / (r[q] + DUMMY_SMALL);
}

__global__ void __launch_bounds__(TPB) k_plain_source(double *__restrict__ rp, const double *__restrict__ g,
                                                      int ni, int i0, int i1, int j0, int j1, double s1,
                                                      double s2) {
#pragma clang fp contract(off)
  const int bw = i1 - i0 + 1, bh = j1 - j0 + 1;
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q >= bw * bh) return;
  const int i = i0 + q % bw, j = j0 + q / bw;
  rp[(size_t)j * ni + i] += s1 * g[q] * s2;
}

// apply_bufferZone_Freund_ambient (source.cpp:174-201), one zone, all nv
// vars (the passive scalars too); a harmonic zone
// (apply_bufferZone_Freund_harmonicWave, source.cpp:264-320) relaxes s' to
// the entropy wave he * hg (hg null: he) and the other variables to 0, the
// data of bc_dirichlet_harmonicwave for LEE.
__global__ void __launch_bounds__(TPB) k_plain_sponge(double *__restrict__ rhs, const double *__restrict__ y,
                                                      const double *__restrict__ sig, int ni, int n, int i0,
                                                      int nzi, int j0, int nzj, int harm, double he,
                                                      const double *__restrict__ hg, int nv) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q >= nzi * nzj) return;
  const int p = (j0 + q / nzi) * ni + (i0 + q % nzi);
  const double s = sig[q];
  for (int v = 0; v < nv; v++) {
    const double target = harm && v == IS ? (hg ? he * hg[q] : he) : 0.0;
    const double relax = -s * (y[(size_t)v * n + p] - target);
    rhs[(size_t)v * n + p] += relax;
  }
}

// RungeKutta::integrate update loop (temporal:208-230)
__global__ void __launch_bounds__(TPB) k_plain_rk(double *__restrict__ rhs, double *__restrict__ ynext,
                                                  double *__restrict__ interim, const double *__restrict__ ycur,
                                                  const int8_t *__restrict__ iblank, int n, double h, double bi,
                                                  double a, int last, int nv) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  if (!iblank || iblank[p] != BLANKED) {
    for (int v = 0; v < nv; v++) {
      const size_t q = (size_t)v * n + p;
      rhs[q] *= h;
      ynext[q] += bi * rhs[q];
      if (!last) interim[q] = ycur[q] + a * rhs[q];
    }
  } else {
    for (int v = 0; v < nv; v++) {
      const size_t q = (size_t)v * n + p;
      rhs[q] = 0.0;
      ynext[q] = 0.0;
      interim[q] = 0.0;
    }
  }
}

// compute_auxiliary_variables_linear_Euler (state.cpp:701-736), rho' only:
// rho' = (p' gamma^-1 p_bar^-1 - s') rho_bar
__global__ void __launch_bounds__(TPB) k_aux_rho(const double *__restrict__ y, const double *__restrict__ mean,
                                                 const double *__restrict__ rhob, double *__restrict__ rho,
                                                 double gi, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  const double pbi = 1.0 / mean[(size_t)IP * n + p];
  double r = y[(size_t)IP * n + p] * gi * pbi - y[(size_t)IS * n + p];
  r *= rhob[p];
  rho[p] = r;
}

// filter blend: y = (1 - blend) func0 + blend func
__global__ void __launch_bounds__(TPB) k_blend(double *__restrict__ y, const double *__restrict__ f0,
                                               const double *__restrict__ f, double rem, double blend, int n) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p < n) y[p] = rem * f0[p] + blend * f[p];
}

// temporal::iblank_correction (temporal:380-393): holes -> sol_ref = 0
__global__ void __launch_bounds__(TPB) k_iblank_reset(double *__restrict__ y, const int8_t *__restrict__ iblank,
                                                      int n, int nv) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= n || iblank[p] != BLANKED) return;
  for (int v = 0; v < nv; v++) y[(size_t)v * n + p] = 0.0;
}

// ------------------------------------------------------ boundary conditions

struct BcArgs {
  int type, dir, end;
  int i0, i1, j0, j1;   // face box (collapsed along dir)
  int ni, nj;
  double neu[9];        // closure row 0 of the FD operator (sb entries)
  int sb;
  double he;            // dirichlet_harmonicwave: entropy factor e(t) at the BC time
  const double *hg;     // and g per face point (null: 1)
  int nsc;              // passive scalars (variables NV .. NV + nsc - 1)
};

// bc::apply_BC one face (boundary_condition.cpp:46-209) with bc_neumann
// (:348-444), bc_wall_slip_kinematic (:448-494), dirichlet_allzero and the
// Cartesian centerline conditions.  Arithmetic in FP64 for both precisions.
// Face point q (of the face box, i fastest) of one BC face.
template <typename T>
__device__ __forceinline__ void bc_face_point(T *__restrict__ y, const double *__restrict__ met,
                                              const double *__restrict__ jac, const int8_t *__restrict__ iblank,
                                              const BcArgs &a, T *__restrict__ rho, const double *__restrict__ mean,
                                              const double *__restrict__ rhob, double gi, int q) {
#pragma clang fp contract(off)
  const int fw = a.i1 - a.i0 + 1;
  const int i = a.i0 + q % fw, j = a.j0 + q / fw;
  const int n = a.ni * a.nj;
  const int p = j * a.ni + i;
  if (iblank && iblank[p] == BLANKED) return;
  const int mf = a.dir == 0 ? 1 : a.ni;
  const int c = a.dir == 0 ? i : j, len = a.dir == 0 ? a.ni : a.nj;
  double data[NV];
  for (int v = 0; v < NV; v++) data[v] = 1.0e9 * (double)(a.dir + 1);
  if (a.type == LEE_BC_DIRICHLET_ALLZERO)
    for (int v = 0; v < NV; v++) data[v] = 0.0;
  if (a.type == LEE_BC_DIRICHLET_HARMONICWAVE) {  // WAVE_ENTROPY: s' = e(t) g, the rest 0 (bc:250-334)
    for (int v = 0; v < NV; v++) data[v] = 0.0;
    data[IS] = a.hg ? a.he * a.hg[q] : a.he;
  }
  const bool neumann = a.type == LEE_BC_NEUMANN || a.type == LEE_BC_WALL_SLIP_KINEMATIC ||
                       (a.type >= LEE_BC_CENTERLINE_CART_NORM2X && a.type <= LEE_BC_CENTERLINE_AXISYM);
  if (neumann) {
    // every neighbour value is requested first (clamped addresses), then the
    // closure row is solved in the reference's order with the skipped
    // entries masked: one memory latency instead of one per entry
    const int sg = a.end == 0 ? 1 : -1;
    T nb[NV][8];
#pragma unroll
    for (int s = 1; s <= 8; s++) {
      const int cs = c + sg * s;
      const int q = p + sg * mf * (cs >= 0 && cs < len ? s : 0);
#pragma unroll
      for (int v = 0; v < NV; v++) nb[v][s - 1] = y[(size_t)v * n + q];
    }
    for (int v = 0; v < NV; v++) {
      if (a.type == LEE_BC_WALL_SLIP_KINEMATIC && v >= IUX && v <= IUZ) continue;
      double d = 0.0;
#pragma unroll
      for (int s = 1; s <= 8; s++) {
        const int cs = c + sg * s;
        const bool use = s < a.sb && cs >= 0 && cs < len;
        const double t1 = a.end == 0 ? d - a.neu[s] * (double)nb[v][s - 1] : d + a.neu[s] * (double)nb[v][s - 1];
        d = use ? t1 : d;
      }
      d /= a.end == 0 ? a.neu[0] : -a.neu[0];
      data[v] = d;
    }
  }
  if (a.type == LEE_BC_WALL_SLIP_KINEMATIC) {
    const double m00 = met[p], m01 = met[(size_t)n + p], m10 = met[(size_t)2 * n + p], m11 = met[(size_t)3 * n + p];
    const double u = (double)y[(size_t)IUX * n + p], v = (double)y[(size_t)IUY * n + p];
    double k0 = 0.0, k1 = 0.0;
    k0 += m00 * u;
    k0 += m01 * v;
    k1 += m10 * u;
    k1 += m11 * v;
    if (a.dir == 0) k0 = 0.0; else k1 = 0.0;
    // metricsInverse (metrics.cpp:375-381): [X][XI] = eta_y, [X][ETA] = -xi_y,
    // [Y][XI] = -eta_x, [Y][ETA] = xi_x
    const double J = jac[p];
    double ux = 0.0, uy = 0.0;
    ux += m11 * k0;
    ux += -m01 * k1;
    ux *= J;
    uy += -m10 * k0;
    uy += m00 * k1;
    uy *= J;
    data[IUX] = ux;
    data[IUY] = uy;
    data[IUZ] = (double)y[(size_t)IUZ * n + p];
  }
  if (a.type >= LEE_BC_CENTERLINE_CART_NORM2X && a.type <= LEE_BC_CENTERLINE_CART_NORM2Z)
    data[IUX + (a.type - LEE_BC_CENTERLINE_CART_NORM2X)] = 0.0;
  if (a.type == LEE_BC_CENTERLINE_AXISYM) data[IUY] = 0.0;  // u_r on the axis (bc:498-545, IVAR_UR)
  // passive scalars: Neumann data at Neumann, wall-slip (its non-velocity
  // variables, boundary_condition.cpp:122-136) and centerline faces; 0 at
  // dirichlet_allzero and harmonic-wave faces, whose data start at 0
  for (int z = 0; z < a.nsc; z++) {
    const size_t vo = (size_t)(NV + z) * n;
    double d = 0.0;
    if (neumann) {
      const int sg = a.end == 0 ? 1 : -1;
      T nbz[8];
#pragma unroll
      for (int s = 1; s <= 8; s++) {
        const int cs = c + sg * s;
        nbz[s - 1] = y[vo + p + sg * mf * (cs >= 0 && cs < len ? s : 0)];
      }
#pragma unroll
      for (int s = 1; s <= 8; s++) {
        const int cs = c + sg * s;
        const bool use = s < a.sb && cs >= 0 && cs < len;
        const double t1 = a.end == 0 ? d - a.neu[s] * (double)nbz[s - 1] : d + a.neu[s] * (double)nbz[s - 1];
        d = use ? t1 : d;
      }
      d /= a.end == 0 ? a.neu[0] : -a.neu[0];
    }
    y[vo + p] = (T)d;
  }
  for (int v = 0; v < NV; v++) y[(size_t)v * n + p] = (T)data[v];
  if (rho) {
    const double pbi = 1.0 / mean[(size_t)IP * n + p];
    double r = data[IP] * gi * pbi - data[IS];
    r *= rhob[p];
    rho[p] = (T)r;
  }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_bc_face(T *__restrict__ y, const double *__restrict__ met,
                                                 const double *__restrict__ jac,
                                                 const int8_t *__restrict__ iblank, BcArgs a,
                                                 T *__restrict__ rho, const double *__restrict__ mean,
                                                 const double *__restrict__ rhob, double gi) {
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q >= (a.i1 - a.i0 + 1) * (a.j1 - a.j0 + 1)) return;
  bc_face_point<T>(y, met, jac, iblank, a, rho, mean, rhob, gi, q);
}

// ---------------------------------------------------------------- overset

template <typename T>
struct OvTables {
  int nv;  // variables per point (NV and the passive scalars)
  T *y[MAX_OV_BLOCKS];
  int ni[MAX_OV_BLOCKS], nn[MAX_OV_BLOCKS];
  T *rho[MAX_OV_BLOCKS];
  const double *mean[MAX_OV_BLOCKS];
  const double *rhob[MAX_OV_BLOCKS];
};

// compute_weighted_sum (overset.cpp:886-974): every receiver's weighted
// donor sum into buf before any receiver is written.
template <typename T>
__global__ void __launch_bounds__(TPB) k_ov_gather(OvTables<T> t, const int *__restrict__ db,
                                                   const int *__restrict__ dl, const int *__restrict__ w,
                                                   const double *__restrict__ coef, int stride,
                                                   double *__restrict__ buf, int nrecv) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * TPB + threadIdx.x;
  if (r >= nrecv) return;
  const int b = db[r];
  const T *y = t.y[b];
  const int ni = t.ni[b], n = t.nn[b];
  const int w1 = w[3 * b], w2 = w[3 * b + 1];
  const double *c = coef + (size_t)r * stride;
  if (w1 == 5 && w2 == 5 && w[3 * b + 2] == 1) {
    // ORDER 4 in 2-D: the 25 donors of one variable are requested together
    // (clamped addresses, skipped donors masked), then summed in the
    // reference's order; one memory latency per variable instead of one
    // per donor
    const int d0 = dl[r];
    double cc[25];
    int li[25];
    bool ok[25];
#pragma unroll
    for (int q = 0; q < 25; q++) {
      const int l1 = d0 + (q % 5) + (q / 5) * ni;
      ok[q] = l1 >= 0 && l1 < n;
      li[q] = ok[q] ? l1 : 0;
      cc[q] = c[q];
    }
    for (int v = 0; v < t.nv; v++) {
      T yv[25];
#pragma unroll
      for (int q = 0; q < 25; q++) yv[q] = y[(size_t)v * n + li[q]];
      double a = 0.0;
#pragma unroll
      for (int q = 0; q < 25; q++) {
        const double t1 = a + cc[q] * (double)yv[q];
        a = ok[q] ? t1 : a;
      }
      buf[(size_t)r * t.nv + v] = a;
    }
    return;
  }
  double acc[MAXV];
  for (int v = 0; v < MAXV; v++) acc[v] = 0.0;
  int s = 0;
  for (int b2 = 0; b2 < w2; b2++)
    for (int b1 = 0; b1 < w1; b1++) {
      const int l1 = dl[r] + b1 + b2 * ni;
      if (l1 >= 0 && l1 < n)
        for (int v = 0; v < t.nv; v++) acc[v] += c[s] * (double)y[(size_t)v * n + l1];
      s++;
    }
  for (int v = 0; v < t.nv; v++) buf[(size_t)r * t.nv + v] = acc[v];
}

// compute_weighted_sum for 5 x 5 stencils (overset.cpp:886-974), receiver r
// and variable v: the 25 donor values and coefficients (receiver-minor, so
// coalesced across receivers) are requested together, then summed in the
// reference's order with the donors outside the block skipped.
template <typename T>
__device__ __forceinline__ double ov_sum5(const OvTables<T> &t, const int *__restrict__ db,
                                          const int *__restrict__ dl, const double *__restrict__ coef5, int nrecv,
                                          int r, int v) {
#pragma clang fp contract(off)
  const int b = db[r];
  const T *y = t.y[b] + (size_t)v * t.nn[b];
  const int ni = t.ni[b], n = t.nn[b], d0 = dl[r];
  double cc[25];
  T yv[25];
  bool ok[25];
#pragma unroll
  for (int q = 0; q < 25; q++) {
    const int l1 = d0 + (q % 5) + (q / 5) * ni;
    ok[q] = l1 >= 0 && l1 < n;
    cc[q] = coef5[(size_t)q * nrecv + r];
    yv[q] = y[ok[q] ? l1 : 0];
  }
  double a = 0.0;
#pragma unroll
  for (int q = 0; q < 25; q++) {
    const double t1 = a + cc[q] * (double)yv[q];
    a = ok[q] ? t1 : a;
  }
  return a;
}

// The gather of the two-phase update: one thread per (receiver, variable),
// every sum into buf before any receiver is written.
template <typename T>
__global__ void __launch_bounds__(TPB) k_ov_gather5(OvTables<T> t, const int *__restrict__ db,
                                                    const int *__restrict__ dl, const double *__restrict__ coef5,
                                                    double *__restrict__ buf, int nrecv) {
  const int gid = blockIdx.x * TPB + threadIdx.x;
  if (gid >= t.nv * nrecv) return;
  const int v = gid / nrecv, r = gid - v * nrecv;
  buf[(size_t)r * t.nv + v] = ov_sum5(t, db, dl, coef5, nrecv, r, v);
}

// update_each_grid (overset.cpp:1241-1278); optionally refreshes rho'.
template <typename T>
__device__ __forceinline__ void ov_scatter_one(const OvTables<T> &t, const int *__restrict__ rb,
                                               const int *__restrict__ rl, const double *__restrict__ buf,
                                               int with_rho, double gi, int r) {
#pragma clang fp contract(off)
  const int b = rb[r];
  const int n = t.nn[b], p = rl[r];
  if (p < 0 || p >= n) return;
  for (int v = 0; v < t.nv; v++) t.y[b][(size_t)v * n + p] = (T)buf[(size_t)r * t.nv + v];
  if (with_rho && t.rho[b]) {
    const double pbi = 1.0 / t.mean[b][(size_t)IP * n + p];
    double q = buf[(size_t)r * t.nv + IP] * gi * pbi - buf[(size_t)r * t.nv + IS];
    q *= t.rhob[b][p];
    t.rho[b][p] = (T)q;
  }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_ov_scatter(OvTables<T> t, const int *__restrict__ rb,
                                                    const int *__restrict__ rl, const double *__restrict__ buf,
                                                    int nrecv, int with_rho, double gi) {
  const int r = blockIdx.x * TPB + threadIdx.x;
  if (r >= nrecv) return;
  ov_scatter_one<T>(t, rb, rl, buf, with_rho, gi, r);
}

// ------------------------------------------------------------------ fused

// INNER: the tile and its halo lie inside a block without holes (interior
// stencils only, no bounds tests); stage_tile below picks the instance.
// W: the derivative operator, OA4 (4, compile-time rows, halo 2) or FDO11P
// (11, rows from a.fdw, halo 6).
// The global loads of one stage tile, held in registers: the tile values
// with their halo (rows ty + 4k, k < KR; columns tx and, for tx < 2H,
// 64 + tx) and the per-point operands of rows ty + 4k (k < R).
template <typename T, int R, int W>
struct StageRegs {
  static constexpr int H = StageOp<W>::H, KR = (4 * R + 2 * H) / 4;
  T ta[4][KR], tb[4][KR];
  double pm[R][4], pj[R];
  T pacc[R][4], pyn[R][4];
};

// (1)-(2): the loads of a tile.  INNER: the tile and its halo lie inside the
// block (no bounds tests); the checked form reads the same values for such
// a tile.
template <typename T, int KIND, int R, bool INNER, int W = 4, typename AT>
__device__ __forceinline__ void stage_load(const AT &a, int tile, StageRegs<T, R, W> &g) {
  constexpr int TY = 4 * R;
  constexpr int H = StageOp<W>::H;
  constexpr int KR = StageRegs<T, R, W>::KR;
  static_assert(TY + 2 * H == 4 * KR, "tile rows and the halo come in fours");
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * ST_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const int VOFF[4] = {IS * n, IUX * n, IUY * n, IP * n};
  constexpr bool inner = INNER;
  const bool xtra = tx < 2 * H;
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int gj = j0 - H + ty + 4 * k;
    const bool rowok = inner || (gj >= 0 && gj < nj);
    const int gia = i0 - H + tx, gib = i0 - H + ST_TX + tx;
    const bool oka = inner || (rowok && gia >= 0 && gia < ni), okb = xtra && (inner || (rowok && gib < ni));
    const int qa = gj * ni + gia, qb = gj * ni + gib;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      g.ta[v][k] = oka ? a.yin[VOFF[v] + qa] : (T)0;
      g.tb[v][k] = okb ? a.yin[VOFF[v] + qb] : (T)0;
    }
  }
  const int gi = i0 + tx;
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int gj = j0 + ty + 4 * k;
    const bool ok = inner || (gi < ni && gj < nj);
    const int p = ok ? gj * ni + gi : 0;
#pragma unroll
    for (int q = 0; q < 4; q++) g.pm[k][q] = ok ? a.met[(size_t)q * n + p] : 0.0;
    g.pj[k] = ok ? a.jac[p] : 0.0;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      g.pacc[k][v] = (KIND != STAGE_FIRST && ok) ? a.acc[VOFF[v] + p] : (T)0;
      g.pyn[k][v] = (KIND == STAGE_MID && ok) ? a.yn[VOFF[v] + p] : (T)0;
    }
  }
}

// (3)-(4): the tile's values through LDS, the derivatives, right-hand side,
// source, sponges and RK update of its points, from the loaded registers.
template <typename T, bool MEANF, bool HOLES, int KIND, int R, bool INNER, int W = 4, typename AT>
__device__ __forceinline__ void stage_finish(const AT &a, const StepT<T> &sc, int tile, const StageRegs<T, R, W> &g,
                                             StageLdsW<T, R, W> &s) {
  // every fusion is written out (fmx): the interior and edge instances of
  // this tile, and the batched kernels, round alike
#pragma clang fp contract(off)
  constexpr int TY = 4 * R;
  constexpr int H = StageOp<W>::H, NBC = StageOp<W>::NB;
  constexpr int KR = StageRegs<T, R, W>::KR;
  constexpr int RS = ST_TX + 2 * H;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * ST_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const int VOFF[4] = {IS * n, IUX * n, IUY * n, IP * n};
  constexpr bool inner = INNER;
  const bool xtra = tx < 2 * H;
  const int gi = i0 + tx;
  const auto &ta = g.ta;
  const auto &tb = g.tb;
  const auto &pm = g.pm;
  const auto &pj = g.pj;
  const auto &pacc = g.pacc;
  const auto &pyn = g.pyn;
  // (3) tile to LDS
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int r = ty + 4 * k;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      s[v][r][tx] = ta[v][k];
      if (xtra) s[v][r][ST_TX + tx] = tb[v][k];
    }
  }
  __syncthreads();

  // (4) the R points
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int rr = ty + 4 * k;
    const int gj = j0 + rr;
    if (!inner && (gi >= ni || gj >= nj)) continue;
    const int p = gj * ni + gi;
    const int lj = rr + H, li = tx + H;
    int cx = CODE_INTERIOR, cy = CODE_INTERIOR;
    if (HOLES) {
      cx = a.code[p];
      cy = a.code[n + p];
    } else if (!inner) {
      cx = default_code(gi, ni, NBC);
      cy = default_code(gj, nj, NBC);
    }
    T dx[4], dy[4], yc[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const T *sp = &s[v][lj][li];
      yc[v] = sp[0];
      if constexpr (W == 4) {
        dx[v] = fd_row<T>(sp, 1, cx);
        dy[v] = fd_row<T>(sp, RS, cy);
      } else {
        dx[v] = wide_row<T>(sp, 1, cx, a.fdw);
        dy[v] = wide_row<T>(sp, RS, cy, a.fdw);
      }
    }
    // chain rule with FP64 metrics (spatial:1810-1820)
    const double mxx = pm[k][0], mxy = pm[k][1], myx = pm[k][2], myy = pm[k][3];
    const double J = pj[k];
    T fx[4], fy[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      fx[v] = (T)(fmx(mxx, (double)dx[v], myx * (double)dy[v]) * J);
      fy[v] = (T)(fmx(mxy, (double)dx[v], myy * (double)dy[v]) * J);
    }
    // LEE right-hand side (spatial:2053-2130), reference term order
    const T sv = yc[0], uv = yc[1], vv = yc[2], pv = yc[3];
    T rs, ru, rv, rp;
    if (!MEANF) {
      const T u0 = (T)a.u0, v0 = (T)a.v0, gp0 = (T)a.gp0;
      rs = -(u0 * fx[0]);
      rs = fmx(-v0, fy[0], rs);
      ru = -(u0 * fx[1]);
      ru = fmx(-v0, fy[1], ru);
      rv = -(u0 * fx[2]);
      rv = fmx(-v0, fy[2], rv);
      if (sizeof(T) == 8) {
        ru -= fx[3] / (T)a.rhob0;
        rv -= fy[3] / (T)a.rhob0;
      } else {
        const T irb = (T)(1.0 / a.rhob0);
        ru = fmx(-fx[3], irb, ru);
        rv = fmx(-fy[3], irb, rv);
      }
      rp = -(u0 * fx[3]);
      rp = fmx(-gp0, fx[1], rp);
      rp = fmx(-v0, fy[3], rp);
      rp = fmx(-gp0, fy[2], rp);
    } else {
      // per-point mean coefficients (MC_* order): u, v, gamma p, 1/rho,
      // the eight mean gradients, rho gi / p and rho (all at the point)
      const T *mc = a.mc;
      const T u0 = mc[MC_U * n + p], v0 = mc[MC_V * n + p], gp = mc[MC_GP * n + p], irb = mc[MC_IRB * n + p];
      const T gsx = mc[MC_GSX * n + p], gsy = mc[MC_GSY * n + p];
      const T gux = mc[MC_GUX * n + p], guy = mc[MC_GUY * n + p];
      const T gvx = mc[MC_GVX * n + p], gvy = mc[MC_GVY * n + p];
      const T gpx = mc[MC_GPX * n + p], gpy = mc[MC_GPY * n + p];
      T rho;
      if (KIND == STAGE_FIRST) rho = a.rho_in[p];
      else rho = fmx(pv, mc[MC_RP * n + p], -(sv * mc[MC_RB * n + p]));
      const T g = (T)a.gamma;
      const T ru_fac = rho * irb;
      const T uf = fmx(ru_fac, u0, uv), vf = fmx(ru_fac, v0, vv), gpv = g * pv;
      rs = -(u0 * fx[0]);
      rs = fmx(-uv, gsx, rs);
      rs = fmx(-v0, fy[0], rs);
      rs = fmx(-vv, gsy, rs);
      ru = -(u0 * fx[1]);
      ru = fmx(-uf, gux, ru);
      ru = fmx(-v0, fy[1], ru);
      ru = fmx(-vf, guy, ru);
      ru = fmx(-fx[3], irb, ru);
      rv = -(u0 * fx[2]);
      rv = fmx(-uf, gvx, rv);
      rv = fmx(-v0, fy[2], rv);
      rv = fmx(-vf, gvy, rv);
      rv = fmx(-fy[3], irb, rv);
      rp = -(u0 * fx[3]);
      rp = fmx(-uv, gpx, rp);
      rp = fmx(-gp, fx[1], rp);
      rp = fmx(-gpv, gux, rp);
      rp = fmx(-v0, fy[3], rp);
      rp = fmx(-vv, gpy, rp);
      rp = fmx(-gp, fy[2], rp);
      rp = fmx(-gpv, gvy, rp);
      // axisymmetry (spatial:2119-2126): -gamma (pbar v' + p' vbar) / (r + 1e-99)
      if (a.axi) rp -= fmx(gp, vv, g * v0 * pv) / mc[MC_R * n + p];
    }
    // physical source (source.cpp:22-107), pressure equation
    if (gi >= a.si0 && gi <= a.si1 && gj >= a.sj0 && gj <= a.sj1) {
      const double gval = a.g[(gj - a.sj0) * (a.si1 - a.si0 + 1) + (gi - a.si0)];
      rp += (T)(sc.s1 * gval * sc.s2);
    }
    // Freund sponges, zone by zone (source.cpp:136-201), sol_ref = 0
#pragma unroll 1
    for (int z = 0; z < a.nz; z++) {
      const auto &zz = a.z[z];
      if (gi >= zz.i0 && gi <= zz.i1 && gj >= zz.j0 && gj <= zz.j1) {
        const int zp = (gj - zz.j0) * zz.nzi + (gi - zz.i0);
        const T sg = zz.sig[zp];
        // a harmonic-wave zone relaxes s' to e(t) g (source.cpp:264-320)
        const T ds = zz.harm ? sv - (T)(zz.hg ? a.he * zz.hg[zp] : a.he) : sv;
        rs = fmx(-sg, ds, rs);
        ru = fmx(-sg, uv, ru);
        rv = fmx(-sg, vv, rv);
        rp = fmx(-sg, pv, rp);
      }
    }
    // RK stage update (temporal:208-230)
    T r[4] = {rs, ru, rv, rp};
    const bool hole = HOLES && a.iblank[p] == BLANKED;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const size_t q = (size_t)VOFF[v] + p;
      T accv = 0, outv = 0;
      if (!hole) {
        const T rh = r[v] * sc.h;
        const T acc_in = KIND == STAGE_FIRST ? yc[v] : pacc[k][v];
        accv = fmx(sc.b, rh, acc_in);
        if (KIND != STAGE_LAST) {
          const T yn = KIND == STAGE_FIRST ? yc[v] : pyn[k][v];
          outv = fmx(sc.a, rh, yn);
        }
      }
      a.acc[q] = accv;
      if (KIND != STAGE_LAST) a.yout[q] = outv;
      r[v] = accv;
    }
    if (MEANF && KIND == STAGE_LAST) a.rho_out[p] = fmx(r[3], a.mc[MC_RP * n + p], -(r[0] * a.mc[MC_RB * n + p]));
  }
}

template <typename T, bool MEANF, bool HOLES, int KIND, int R, bool INNER, int W = 4, typename AT>
__device__ __forceinline__ void stage_tile_k(const AT &a, const StepT<T> &sc, int tile, StageLdsW<T, R, W> &s) {
  StageRegs<T, R, W> g;
  stage_load<T, KIND, R, INNER, W>(a, tile, g);
  stage_finish<T, MEANF, HOLES, KIND, R, INNER, W>(a, sc, tile, g, s);
}

template <typename T, bool MEANF, bool HOLES, int KIND, int R, int W = 4, typename AT>
__device__ __forceinline__ void stage_tile(const AT &a, const StepT<T> &sc, int tile, StageLdsW<T, R, W> &s) {
  constexpr int TY = 4 * R, H = StageOp<W>::H;
  const int i0 = (tile % a.ntx) * ST_TX, j0 = (tile / a.ntx) * TY;
  if constexpr (!HOLES) {
    if (i0 >= H && i0 + ST_TX + H <= a.ni && j0 >= H && j0 + TY + H <= a.nj) {
      stage_tile_k<T, MEANF, HOLES, KIND, R, true, W>(a, sc, tile, s);
      return;
    }
  } else {
    // a tile of a block with holes whose points all take the interior
    // stencil (no hole, no closure row) runs the interior instance
    if (a.clean && box_clean(a.clean, a.ni, a.nj, i0, i0 + ST_TX, j0, j0 + TY)) {
      stage_tile_k<T, MEANF, false, KIND, R, true, W>(a, sc, tile, s);
      return;
    }
  }
  stage_tile_k<T, MEANF, HOLES, KIND, R, false, W>(a, sc, tile, s);
}

template <typename T, bool MEANF, bool HOLES, int KIND, int R, int W = 4>
__global__ void __launch_bounds__(TPB) k_fused_stage(StageArgs<T> a) {
  __shared__ StageLdsW<T, R, W> s;
  const int tile = xcd_tile(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  const StepT<T> sc{a.h, a.b, a.a, a.s1, a.s2};
  stage_tile<T, MEANF, HOLES, KIND, R, W>(a, sc, tile, s);
}

// Two blocks of one handle in one launch (an overset pair such as a
// background grid and a body grid): block 0's tiles, then block 1's, as one
// XCD-mapped list, each tile with its block's operands.  Every point runs
// the single-block arithmetic, so the results are those of two launches;
// the small block no longer costs a launch of its own.
template <typename T, bool MEANF, bool HOLES, int KIND, int R>
__global__ void __launch_bounds__(TPB) k_fused_stage_pair(StageArgs<T> a0, StageArgs<T> a1) {
  __shared__ StageLdsW<T, R, 4> s;
  const int total = a0.ntiles + a1.ntiles;
  const int tile = xcd_tile(blockIdx.x, (total + 7) / 8);
  if (tile >= total) return;
  const StepT<T> sc{a0.h, a0.b, a0.a, a0.s1, a0.s2};
  if (tile < a0.ntiles) stage_tile<T, MEANF, HOLES, KIND, R, 4>(a0, sc, tile, s);
  else stage_tile<T, MEANF, HOLES, KIND, R, 4>(a1, sc, tile - a0.ntiles, s);
}

// Batched ensemble: the tiles of all solves form one list; map[2t] is the
// solve of list entry t and map[2t + 1] its tile within the solve.  The
// list is split into eighths of `chunk` entries, one per XCD: workgroup
// (or virtual index) v maps to list position (v & 7) * chunk + (v >> 3),
// so with a grid that is a multiple of 8 each XCD works through one
// contiguous eighth of the list.  Per-solve operands come from `args`, the
// per-step scalars from `sched`, indexed by the step counter that the
// step's last kernel advances.
template <typename T>
struct BatchStageArgs {
  const StageArgs<T> *args;      // nsolve entries for this stage
  const int *map;                // (solve, tile) per list entry
  const BatchSched *sched;       // [step][solve][stage]
  const int *counter;            // step index within the schedule
  int nsolve, nstage, stage, total, chunk;
  T b, a;
};

// The operand tables are read through the constant address space: they do
// not change during a launch, so their loads are scalar and the pointers
// loaded from them are known to address global memory (no flat accesses).
// One workgroup per list entry (the queue argument is unused: this kernel
// shares its signature with the persistent form below).
template <typename T, bool MEANF, int KIND, int R>
__global__ void __launch_bounds__(TPB)
    k_batch_stage(BatchStageArgs<T> ba, const LEE_CONST_AS StageArgs<T> *args, const LEE_CONST_AS int *map,
                  const LEE_CONST_AS BatchSched *sched, const int *__restrict__ counter, int *) {
  __shared__ StageLds<T, R> s;
  const int t = (blockIdx.x & 7) * ba.chunk + (blockIdx.x >> 3);
  if (t >= ba.total) return;
  // device-scope load: the counter was written by the previous step's filter launch
  const int step = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int k = map[2 * t], local = map[2 * t + 1];
  const auto &q = sched[((size_t)step * ba.nsolve + k) * ba.nstage + ba.stage];
  const StepT<T> sc{(T)q.dt, ba.b, ba.a, q.s1, q.s2};
  stage_tile<T, MEANF, false, KIND, R>(args[k], sc, local, s);
}

// Persistent form of the batched stage (a grid smaller than the list).
// Workgroup b works on eighth b & 7 of the list (the hardware places
// workgroup b on XCD b mod 8, so each XCD keeps one contiguous eighth),
// one of the W = grid / 8 workgroups of that eighth.  The first S entries
// of the eighth (7/8 of them, whole rounds of W) are dealt out statically,
// entry w + k W to workgroup w; the rest are drawn QCH at a time from an
// atomic counter, so that the workgroups whose static share held cheap
// tiles (interior, outside sponges) take more of the remainder and the
// launch ends with the list instead of with the slowest fixed share.
// queue[QSTRIDE x] counts the draws from eighth x and queue[QSTRIDE x + 1]
// the workgroups of eighth x that have finished; the last of those resets
// both, so every launch starts from zero (each eighth has its own cache
// line).  A workgroup stops once its entry passes the eighth's end, and
// every draw advances the counter, so the loop ends.
constexpr int QCH = 2, QSTRIDE = 32;

template <typename T, bool MEANF, int KIND, int R>
__global__ void __launch_bounds__(TPB)
    k_batch_stage_q(BatchStageArgs<T> ba, const LEE_CONST_AS StageArgs<T> *args, const LEE_CONST_AS int *map,
                    const LEE_CONST_AS BatchSched *sched, const int *__restrict__ counter, int *__restrict__ queue) {
  __shared__ StageLds<T, R> s;
  __shared__ int drawn[2];
  const int step = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int x = blockIdx.x & 7, lo = x * ba.chunk;
  const int n = min(ba.chunk, ba.total - lo);  // entries of this eighth (may be <= 0)
  const int W = (int)(gridDim.x >> 3), w = (int)(blockIdx.x >> 3);
  const int S = n > 0 ? (n * 7 / 8) / W * W : 0;
  int *const qx = queue + QSTRIDE * x;
  const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
  // draw d goes through slot d & 1: every thread reads it right after draw
  // d's barrier, so before draw d + 1's barrier, and the lead thread writes
  // the slot again only at draw d + 2, after that barrier
  int nd = 0;
  auto draw = [&]() {
    const int slot = nd++ & 1;
    if (lead) drawn[slot] = __hip_atomic_fetch_add(qx, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return S + QCH * drawn[slot];
  };
  auto advance = [&](int e) {
    if (e < S) return e + W < S ? e + W : draw();
    if ((e - S) % QCH != QCH - 1) return e + 1;
    return draw();
  };
  // the loads take the bounds-checked form, which reads the same values for
  // an interior tile; the finish takes the tile's own instance
  auto finish = [&](int e, const StageRegs<T, R, 4> &g) {
    const int t = lo + e;
    const int k = map[2 * t], local = map[2 * t + 1];
    const auto &q = sched[((size_t)step * ba.nsolve + k) * ba.nstage + ba.stage];
    const StepT<T> sc{(T)q.dt, ba.b, ba.a, q.s1, q.s2};
    const int i0 = (local % args[k].ntx) * ST_TX, j0 = (local / args[k].ntx) * (4 * R);
    if (i0 >= ST_H && i0 + ST_TX + ST_H <= args[k].ni && j0 >= ST_H && j0 + 4 * R + ST_H <= args[k].nj)
      stage_finish<T, MEANF, false, KIND, R, true>(args[k], sc, local, g, s);
    else
      stage_finish<T, MEANF, false, KIND, R, false>(args[k], sc, local, g, s);
  };
  int e = w < S ? w : draw();
  while (e < n) {
    const int t = lo + e;
    StageRegs<T, R, 4> g;
    stage_load<T, KIND, R, false>(args[map[2 * t]], map[2 * t + 1], g);
    finish(e, g);
    __syncthreads();  // the LDS tile is rewritten by the next entry
    e = advance(e);
  }
  if (lead) {
    const int done = __hip_atomic_fetch_add(qx + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (done == W - 1) {
      __hip_atomic_store(qx, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(qx + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

template <typename T, bool HOLES, int ORDER, int NVAR, int TY>
__global__ void __launch_bounds__(TPB) k_fused_filter(FilterArgs<T> a) {
  __shared__ FilterLds<T, NVAR, TY, filter_halo<ORDER>()> L;
  const int tile = xcd_tile(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  filter_tile<T, HOLES, ORDER, NVAR, TY>(a, tile, L);
}

// Persistent, software-pipelined form: a grid of G workgroups (G a
// multiple of 8, about one resident wave of workgroups) strides through
// the XCD-permuted tile list; each workgroup issues the global loads of its
// next tile before computing the current one, so the loads overlap the
// LDS sweeps instead of waiting for a new launch round.  Two tiles' values
// stay in registers: up to 256 VGPRs (two waves per SIMD) keep them out of
// scratch.
template <typename T, bool HOLES, int ORDER, int NVAR, int TY>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(1, 2)))
k_fused_filter_p(FilterArgs<T> a) {
  __shared__ FilterLds<T, NVAR, TY> L;
  const int G = gridDim.x, total = 8 * a.chunk;
  auto tile_of = [&](int v) { return (v & 7) * a.chunk + (v >> 3); };
  auto live = [&](int v) { return v < total && tile_of(v) < a.ntiles; };
  // loads are issued for every slot (past the end: the last tile again) so
  // that no branch sits between a tile's loads and its use
  auto ld = [&](int v) { return min(tile_of(v), a.ntiles - 1); };
  // (the scheduling barrier keeps the compiler from sinking the prefetch
  // into the sweeps, where each load would be waited for at once; the copy
  // at the end of an iteration waits for loads that had the whole tile's
  // sweeps to arrive)
  FilterRegs<T, NVAR, TY> R;
  int v = blockIdx.x;
  filter_load<T, HOLES, NVAR, TY>(a, ld(v), R);
  for (; v < total; v += G) {
    FilterRegs<T, NVAR, TY> N;
    filter_load<T, HOLES, NVAR, TY>(a, ld(v + G), N);
    __builtin_amdgcn_sched_barrier(0);
    if (live(v)) filter_compute<T, HOLES, ORDER, NVAR, TY>(a, tile_of(v), L, R);
    __syncthreads();
    R = N;
  }
}

// Column-march form (blocks without holes): one wave per (variable, strip,
// row chunk), four waves per workgroup, no LDS (filter_march).
template <typename T, int ORDER, int L>
__global__ void __launch_bounds__(TPB) k_fused_filter_m(FilterArgs<T> a) {
  filter_march<T, ORDER, L>(a);
}

// Stacked column march (blocks without holes): NW waves per (variable,
// strip, stack of NW R rows), xi rows shared through LDS (filter_stack).
template <typename T, int ORDER, int NW, int R, bool HOLES>
__global__ void __launch_bounds__(64 * NW) k_fused_filter_s(FilterArgs<T> a) {
  __shared__ StackLds<T, NW, R> L;
  filter_stack<T, ORDER, NW, R, HOLES>(a, L);
}

// The stacked filter of two blocks in one launch (see k_fused_stage_pair).
template <typename T, int ORDER, int NW, int R, bool HOLES>
__global__ void __launch_bounds__(64 * NW) k_fused_filter_s_pair(FilterArgs<T> a0, FilterArgs<T> a1) {
  __shared__ StackLds<T, NW, R> L;
  const int total = a0.ntiles + a1.ntiles;
  const int t = xcd_tile(blockIdx.x, (total + 7) / 8);
  if (t >= total) return;
  if (t < a0.ntiles) stack_item<T, ORDER, NW, R, HOLES>(a0, t, L);
  else stack_item<T, ORDER, NW, R, HOLES>(a1, t - a0.ntiles, L);
}

template <typename T>
struct BatchFilterArgs {
  const FilterArgs<T> *args;     // nsolve entries
  const int *map;                // (solve, tile) per list entry
  int *counter;                  // advanced by one at the end of the step
  int nsolve, total, chunk;
};

// The batched filter is the last kernel of a step: one thread advances the
// step counter (no kernel of this launch reads it; the next step's stage
// kernels see the new value).
template <typename T, int ORDER, int TY>
__global__ void __launch_bounds__(TPB)
    k_batch_filter(BatchFilterArgs<T> ba, const LEE_CONST_AS FilterArgs<T> *args, const LEE_CONST_AS int *map) {
  __shared__ FilterLds<T, 4, TY> L;
  if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    const int c = __hip_atomic_load(ba.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(ba.counter, c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // vector store
  }
  for (int v = blockIdx.x; v < 8 * ba.chunk; v += gridDim.x) {
    const int t = (v & 7) * ba.chunk + (v >> 3);
    if (t >= ba.total) continue;
    filter_tile<T, false, ORDER, 4, TY>(args[map[2 * t]], map[2 * t + 1], L);
    __syncthreads();
  }
}


// Batched stacked column march (solves without holes): the list entries are
// (solve, stack) pairs; four waves of R rows per stack.
template <typename T, int ORDER, int R>
__global__ void __launch_bounds__(TPB)
    k_batch_filter_s(BatchFilterArgs<T> ba, const LEE_CONST_AS FilterArgs<T> *args, const LEE_CONST_AS int *map) {
  __shared__ StackLds<T, 4, R> L;
  if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    const int c = __hip_atomic_load(ba.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(ba.counter, c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // vector store
  }
  for (int v = blockIdx.x; v < 8 * ba.chunk; v += gridDim.x) {
    const int t = (v & 7) * ba.chunk + (v >> 3);
    if (t >= ba.total) continue;
    stack_item<T, ORDER, 4, R>(args[map[2 * t]], map[2 * t + 1], L);
    __syncthreads();  // the LDS rows are rewritten by the next stack
  }
}

// Passive w' in 2-D: only the sponge acts on it (the velocity loop of
// spatial:2077-2103 stops at num_dim), then the RK update.
template <typename T>
__global__ void __launch_bounds__(TPB) k_passive_w(StageArgs<T> a, int kind) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  if (p >= n) return;
  const int gi = p % ni, gj = p / ni;
  const size_t q = (size_t)IUZ * n + p;
  const T w = a.yin[q];
  T r = 0;
  for (int z = 0; z < a.nz; z++) {
    const ZoneT<T> &zz = a.z[z];
    if (gi >= zz.i0 && gi <= zz.i1 && gj >= zz.j0 && gj <= zz.j1)
      r += -zz.sig[(gj - zz.j0) * zz.nzi + (gi - zz.i0)] * w;
  }
  if (a.iblank && a.iblank[p] == BLANKED) {
    a.acc[q] = 0;
    if (kind != STAGE_LAST) a.yout[q] = 0;
    return;
  }
  const T rh = r * a.h;
  const T acc_in = kind == STAGE_FIRST ? w : a.acc[q];
  a.acc[q] = acc_in + a.b * rh;
  if (kind != STAGE_LAST) {
    const T yn = kind == STAGE_FIRST ? w : a.yn[q];
    a.yout[q] = yn + a.a * rh;
  }
}

// One RK stage of the passive scalars of a LEE_SCALAR run, one thread per
// (scalar, point), after the stage kernel of the same buffers (it writes
// only the five LEE variables of acc and yout; this kernel only the
// scalars'): compute_RHS_linearized_Euler_scalar (spatial:2134-2158),
//   rhs_Z = -ubar dZ'/dx - u' dZbar/dx - vbar dZ'/dy - v' dZbar/dy
// in the reference's term order, the Freund sponges relaxing Z' to 0
// (source.cpp:174-201, 264-320), the RK update and the hole reset as in
// the stage kernel.  Derivative rows read the stage input in place (the
// closure rows stay inside the block); metrics, J, the mean velocity and
// the scalars' mean gradients are FP64.  W: OA4 (4) or FDO11P (11) rows.
template <typename T, int KIND, int W>
__global__ void __launch_bounds__(TPB) k_scalar_stage(StageArgs<T> a, int nsc) {
#pragma clang fp contract(off)
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= nsc * n) return;
  const int z = t / n, p = t - z * n;
  const int gi = p % ni, gj = p / ni;
  const size_t q = (size_t)(NV + z) * n + p;
  int cx, cy;
  if (a.code) {
    cx = a.code[p];
    cy = a.code[n + p];
  } else {
    cx = default_code(gi, ni, StageOp<W>::NB);
    cy = default_code(gj, nj, StageOp<W>::NB);
  }
  const T *sp = a.yin + q;
  T dx, dy;
  if constexpr (W == 4) {
    dx = fd_row<T>(sp, 1, cx);
    dy = fd_row<T>(sp, ni, cy);
  } else {
    dx = wide_row<T>(sp, 1, cx, a.fdw);
    dy = wide_row<T>(sp, ni, cy, a.fdw);
  }
  const double mxx = a.met[p], mxy = a.met[(size_t)n + p], myx = a.met[(size_t)2 * n + p],
               myy = a.met[(size_t)3 * n + p];
  const double J = a.jac[p];
  const T fx = (T)(fmx(mxx, (double)dx, myx * (double)dy) * J);
  const T fy = (T)(fmx(mxy, (double)dx, myy * (double)dy) * J);
  const T u0 = (T)a.mean[(size_t)IUX * n + p], v0 = (T)a.mean[(size_t)IUY * n + p];
  const T up = a.yin[(size_t)IUX * n + p], vp = a.yin[(size_t)IUY * n + p];
  const T gzx = (T)a.mgrad[(size_t)(2 * (NV + z)) * n + p], gzy = (T)a.mgrad[(size_t)(2 * (NV + z) + 1) * n + p];
  const T zc = sp[0];
  T r = -(u0 * fx);
  r = fmx(-up, gzx, r);
  r = fmx(-v0, fy, r);
  r = fmx(-vp, gzy, r);
  for (int k = 0; k < a.nz; k++) {
    const ZoneT<T> &zz = a.z[k];
    if (gi >= zz.i0 && gi <= zz.i1 && gj >= zz.j0 && gj <= zz.j1)
      r = fmx(-zz.sig[(gj - zz.j0) * zz.nzi + (gi - zz.i0)], zc, r);
  }
  T accv = 0, outv = 0;
  if (!(a.iblank && a.iblank[p] == BLANKED)) {
    const T rh = r * a.h;
    const T acc_in = KIND == STAGE_FIRST ? zc : a.acc[q];
    accv = fmx(a.b, rh, acc_in);
    if (KIND != STAGE_LAST) {
      const T yn = KIND == STAGE_FIRST ? zc : a.yn[q];
      outv = fmx(a.a, rh, yn);
    }
  }
  a.acc[q] = accv;
  if (KIND != STAGE_LAST) a.yout[q] = outv;
}

template <typename S, typename D>
__global__ void __launch_bounds__(TPB) k_cast(const S *__restrict__ a, D *__restrict__ b, size_t n) {
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p < n) b[p] = (D)a[p];
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

}  // namespace

// ------------------------------------------------------------- launchers

hipError_t plain_apply_op(hipStream_t s, const Block &b, const OpTable &op, const int8_t *codes, const double *f,
                          double *df, int dir) {
  hipLaunchKernelGGL(k_plain_op_dir, dim3(nblocks(b.n)), dim3(TPB), 0, s, f, df, codes, b.ni, b.nj, dir, op);
  return hipGetLastError();
}

hipError_t metrics2d(hipStream_t s, Block &b, const OpTable &fd) {
  // x_xi, x_eta, y_xi, y_eta by the FD operator with no holes (metrics are
  // computed before overset marks any, simulation.cpp:207-217)
  double *xx = nullptr;
  hipError_t e = hipMalloc((void **)&xx, sizeof(double) * 4 * (size_t)b.n);
  if (e != hipSuccess) return e;
  for (int ph = 0; ph < 2; ph++)
    for (int cp = 0; cp < 2; cp++) {
      hipLaunchKernelGGL(k_plain_op_dir, dim3(nblocks(b.n)), dim3(TPB), 0, s, b.xyz + (size_t)ph * b.n,
                         xx + (size_t)(2 * ph + cp) * b.n, (const int8_t *)nullptr, b.ni, b.nj, cp, fd);
    }
  hipLaunchKernelGGL(k_metrics2d, dim3(nblocks(b.n)), dim3(TPB), 0, s, xx, b.met, b.jac, b.n);
  e = hipGetLastError();
  hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(xx);
  return e != hipSuccess ? e : e2;
}

hipError_t plain_take_derivative_xyz(hipStream_t s, const Block &b, const OpTable &op, const double *f,
                                     int dir_xyz) {
  hipLaunchKernelGGL(k_plain_deriv_xyz, dim3(nblocks(b.n)), dim3(TPB), 0, s, f, b.dflux, b.met, b.jac,
                     b.has_holes ? b.code_fd : nullptr, b.ni, b.nj, dir_xyz, op);
  return hipGetLastError();
}

hipError_t aux_density(hipStream_t s, const Block &b, double gamma, const double *y, double *rho) {
  hipLaunchKernelGGL(k_aux_rho, dim3(nblocks(b.n)), dim3(TPB), 0, s, y, b.mean, b.rhobar, rho, 1.0 / gamma, b.n);
  return hipGetLastError();
}

// spatial::compute_RHS + compute_RHS_linearized_Euler + add_RHSSourceTerms
// (spatial:1963-1990, 2053-2130; source.cpp:10-18), one kernel per array
// operation of the reference.  rho: the aux density to use; anything other
// than the block's carried array is first filled from y.
hipError_t plain_rhs(hipStream_t s, lee_handle *h, const Block &b, const double *y, double *rhs, double s1,
                     double s2, double *rho, double he) {
  const int n = b.n;
  const dim3 g(nblocks(n)), t(TPB);
  const double gam = h->cfg.gamma;
  auto Y = [&](int v) { return y + (size_t)v * n; };
  auto R = [&](int v) { return rhs + (size_t)v * n; };
  auto M = [&](int v) { return b.mean + (size_t)v * n; };
  auto G = [&](int v, int d) { return b.mgrad + (size_t)(2 * v + d) * n; };
  if (rho != b.rho64) {  // unit call: aux density of y itself, carried rho' untouched
    hipError_t e = aux_density(s, b, gam, y, rho);
    if (e != hipSuccess) return e;
  }
  hipError_t err = hipSuccess;  // first failure of the derivative launches
  auto deriv = [&](const double *f, int d) {
    const hipError_t e = plain_take_derivative_xyz(s, b, h->fd, f, d);
    if (err == hipSuccess) err = e;
  };
  const int nv = nvars(h);
  hipLaunchKernelGGL(k_fill, dim3(nblocks((size_t)nv * n)), t, 0, s, rhs, 0.0, (size_t)nv * n);
  for (int d = 0; d < 2; d++) {
    deriv(Y(IS), d);
    hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(IS), M(IUX + d), b.dflux, n);
    hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(IS), Y(IUX + d), G(IS, d), n);
  }
  for (int vel = 0; vel < 2; vel++) {
    for (int d = 0; d < 2; d++) {
      deriv(Y(IUX + vel), d);
      hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(IUX + vel), M(IUX + d), b.dflux, n);
      hipLaunchKernelGGL(k_sub_velgrad, g, t, 0, s, R(IUX + vel), Y(IUX + d), rho, b.rhobar, M(IUX + d),
                         G(IUX + vel, d), n);
    }
    deriv(Y(IP), vel);
    hipLaunchKernelGGL(k_sub_div, g, t, 0, s, R(IUX + vel), b.dflux, b.rhobar, n);
  }
  for (int d = 0; d < 2; d++) {
    deriv(Y(IP), d);
    hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(IP), M(IUX + d), b.dflux, n);
    hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(IP), Y(IUX + d), G(IP, d), n);
    deriv(Y(IUX + d), d);
    hipLaunchKernelGGL(k_sub_gab, g, t, 0, s, R(IP), gam, M(IP), b.dflux, n);
    hipLaunchKernelGGL(k_sub_gab, g, t, 0, s, R(IP), gam, Y(IP), G(IUX + d, d), n);
  }
  if (h->cfg.axisym)
    hipLaunchKernelGGL(k_plain_axisym, g, t, 0, s, R(IP), gam, M(IP), Y(IUY), Y(IP), M(IUY), b.xyz + n, n);
  // passive scalars, compute_RHS_linearized_Euler_scalar (spatial:2134-2158)
  // after the LEE terms (spatial:1977-1979)
  for (int v = NV; v < nv; v++)
    for (int d = 0; d < 2; d++) {
      deriv(Y(v), d);
      hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(v), M(IUX + d), b.dflux, n);
      hipLaunchKernelGGL(k_sub_ab, g, t, 0, s, R(v), Y(IUX + d), G(v, d), n);
    }
  if (b.src.i1 >= b.src.i0) {
    const size_t cnt = (size_t)(b.src.i1 - b.src.i0 + 1) * (b.src.j1 - b.src.j0 + 1);
    hipLaunchKernelGGL(k_plain_source, dim3(nblocks(cnt)), t, 0, s, R(IP), b.src.g, b.ni, b.src.i0, b.src.i1,
                       b.src.j0, b.src.j1, s1, s2);
  }
  for (const auto &z : b.zones) {
    const size_t cnt = (size_t)z.nz[0] * z.nz[1];
    hipLaunchKernelGGL(k_plain_sponge, dim3(nblocks(cnt)), t, 0, s, rhs, y, z.sigma64, b.ni, n, z.is[0], z.nz[0],
                       z.is[1], z.nz[1], z.harmonic ? 1 : 0, he, z.hg, nv);
  }
  const hipError_t last = hipGetLastError();
  return err != hipSuccess ? err : last;
}

hipError_t plain_rk_update(hipStream_t s, const Block &b, double *rhs, double *ynext, double *interim,
                           const double *ycur, double hdt, double bi, double a, bool last, int nv) {
  hipLaunchKernelGGL(k_plain_rk, dim3(nblocks(b.n)), dim3(TPB), 0, s, rhs, ynext, interim, ycur,
                     b.has_holes ? b.iblank : nullptr, b.n, hdt, bi, a, last ? 1 : 0, nv);
  return hipGetLastError();
}

hipError_t plain_filter(hipStream_t s, lee_handle *h, const Block &b, double *y) {
  const int n = b.n;
  const dim3 g(nblocks(n)), t(TPB);
  const double blend = h->cfg.filter_blend, rem = 1.0 - blend;
  for (int v = 0; v < nvars(h); v++) {
    double *yv = y + (size_t)v * n;
    hipLaunchKernelGGL(k_copy, g, t, 0, s, yv, b.func0, (size_t)n);
    hipLaunchKernelGGL(k_copy, g, t, 0, s, b.func0, b.func, (size_t)n);
    for (int d = 0; d < 2; d++) {
      hipLaunchKernelGGL(k_plain_op_dir, g, t, 0, s, b.func, b.dflux,
                         b.has_holes ? b.code_flt + (size_t)d * n : nullptr, b.ni, b.nj, d, h->flt);
      hipLaunchKernelGGL(k_copy, g, t, 0, s, b.dflux, b.func, (size_t)n);
    }
    hipLaunchKernelGGL(k_blend, g, t, 0, s, yv, b.func0, b.func, rem, blend, n);
  }
  return hipGetLastError();
}

hipError_t plain_iblank_reset(hipStream_t s, const Block &b, double *y, int nv) {
  if (!b.has_holes) return hipSuccess;
  hipLaunchKernelGGL(k_iblank_reset, dim3(nblocks(b.n)), dim3(TPB), 0, s, y, b.iblank, b.n, nv);
  return hipGetLastError();
}

static BcArgs make_bc_args(const lee_handle *h, const Block &b, const BcFace &f, double he) {
  BcArgs a{};
  a.he = he;
  a.hg = f.hg;
  a.type = f.type;
  a.dir = f.dir;
  a.end = f.end;
  a.i0 = f.is[0];
  a.i1 = f.ie[0];
  a.j0 = f.is[1];
  a.j1 = f.ie[1];
  a.ni = b.ni;
  a.nj = b.nj;
  a.sb = h->fd.sb;
  a.nsc = h->cfg.num_scalar;
  for (int q = 0; q < 9; q++) a.neu[q] = h->fd.bnd[0][q];
  return a;
}

hipError_t apply_bcs(hipStream_t s, lee_handle *h, const Block &b, double *y64, float *y32, bool fp32,
                     double *rho64, float *rho32, double he) {
  for (const auto &f : b.faces) {
    const BcArgs a = make_bc_args(h, b, f, he);
    const size_t cnt = (size_t)(a.i1 - a.i0 + 1) * (a.j1 - a.j0 + 1);
    const int8_t *ib = b.has_holes ? b.iblank : nullptr;
    if (fp32)
      hipLaunchKernelGGL(k_bc_face<float>, dim3(nblocks(cnt)), dim3(TPB), 0, s, y32, b.met, b.jac, ib, a, rho32,
                         b.mean, b.rhobar, 1.0 / h->cfg.gamma);
    else
      hipLaunchKernelGGL(k_bc_face<double>, dim3(nblocks(cnt)), dim3(TPB), 0, s, y64, b.met, b.jac, ib, a, rho64,
                         b.mean, b.rhobar, 1.0 / h->cfg.gamma);
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t overset_t(hipStream_t s, lee_handle *h, T *const *ys, T *const *rhos, bool with_rho, double *buf) {
  OvTables<T> t{};
  t.nv = nvars(h);
  for (int ib = 0; ib < h->nblk; ib++) {
    t.y[ib] = ys[ib];
    t.ni[ib] = h->b[ib].ni;
    t.nn[ib] = h->b[ib].n;
    t.rho[ib] = rhos ? rhos[ib] : nullptr;
    t.mean[ib] = h->b[ib].mean;
    t.rhob[ib] = h->b[ib].rhobar;
  }
  const Overset &o = h->ov;
  if (o.coef5)
    hipLaunchKernelGGL(k_ov_gather5<T>, dim3(nblocks((size_t)t.nv * o.nrecv)), dim3(TPB), 0, s, t, o.db, o.dl, o.coef5,
                       buf, o.nrecv);
  else
    hipLaunchKernelGGL(k_ov_gather<T>, dim3(nblocks(o.nrecv)), dim3(TPB), 0, s, t, o.db, o.dl, o.w, o.coef, o.stride,
                       buf, o.nrecv);
  hipLaunchKernelGGL(k_ov_scatter<T>, dim3(nblocks(o.nrecv)), dim3(TPB), 0, s, t, o.rb, o.rl, buf, o.nrecv,
                     with_rho ? 1 : 0, 1.0 / h->cfg.gamma);
  return hipGetLastError();
}

// buf_index selects the state buffer (Y[k]) of every block; buf_index < 0
// means the plain path's arrays (-1 interim, -2 ynp1).
hipError_t overset_interpolate(hipStream_t s, lee_handle *h, bool fp32, int buf_index, bool last_stage) {
  if (h->ov.nrecv == 0) return hipSuccess;
  if (fp32) {
    float *ys[MAX_OV_BLOCKS], *rh[MAX_OV_BLOCKS];
    bool any_rho = false;
    for (int ib = 0; ib < h->nblk; ib++) {
      ys[ib] = h->b[ib].Y32[buf_index];
      rh[ib] = h->b[ib].uniform_mean ? nullptr : h->b[ib].rho32;
      any_rho |= rh[ib] != nullptr;
    }
    return overset_t<float>(s, h, ys, rh, last_stage && any_rho, h->ov.buf);
  }
  double *ys[MAX_OV_BLOCKS], *rh[MAX_OV_BLOCKS];
  bool any_rho = false;
  for (int ib = 0; ib < h->nblk; ib++) {
    if (buf_index >= 0) ys[ib] = h->b[ib].Y64[buf_index];
    else ys[ib] = buf_index == -1 ? h->b[ib].interim : h->b[ib].ynp1;
    rh[ib] = (buf_index >= 0 && !h->b[ib].uniform_mean) ? h->b[ib].rho64 : nullptr;
    any_rho |= rh[ib] != nullptr;
  }
  return overset_t<double>(s, h, ys, rh, last_stage && any_rho, h->ov.buf);
}

hipError_t overset_interpolate_arrays(hipStream_t s, lee_handle *h, double *const *ys, double *buf) {
  if (h->ov.nrecv == 0) return hipSuccess;
  return overset_t<double>(s, h, ys, nullptr, false, buf);
}

// The boundary phase of a fused stage on the stage's target buffers: the
// BC faces of every block (a launch per face), then the overset
// interpolation (gather, then scatter).  One cooperative launch with grid
// barriers in between was measured slower (overset cylinder 1.72e10
// against 2.18e10 FP32 updates/s over 1000 steps): the runtime's
// cooperative launch costs more than the launches it replaces.
hipError_t boundary_phase_fused(hipStream_t s, lee_handle *h, bool fp32, int tgt, bool last, double he) {
  hipError_t e = hipSuccess;
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    if (b.faces.empty()) continue;
    const bool rho = last && !b.uniform_mean;
    e = apply_bcs(s, h, b, b.Y64[tgt], b.Y32[tgt], fp32, rho ? b.rho64 : nullptr, rho ? b.rho32 : nullptr, he);
    if (e != hipSuccess) return e;
  }
  return overset_interpolate(s, h, fp32, tgt, last);
}

template <typename T>
static void fill_zones(StageArgs<T> &a, const Block &b) {
  a.nz = 0;
  for (const auto &z : b.zones) {
    if (a.nz >= MAX_ZONES) break;
    ZoneT<T> &q = a.z[a.nz++];
    q.i0 = z.is[0];
    q.i1 = z.ie[0];
    q.j0 = z.is[1];
    q.j1 = z.ie[1];
    q.nzi = z.nz[0];
    if constexpr (sizeof(T) == 8) q.sig = (const T *)z.sigma64;
    else q.sig = (const T *)z.sigma32;
    q.harm = z.harmonic ? 1 : 0;
    q.hg = z.hg;
  }
}

// Stage kind of RK stage `stage` (stage < 0: right-hand side only).
static int stage_kind(const lee_handle *h, int stage) {
  const int last = h->cfg.rk_stages - 1;
  return stage <= 0 ? (stage == 0 ? STAGE_FIRST : STAGE_MID) : (stage == last ? STAGE_LAST : STAGE_MID);
}

// Operands of one RK stage on block b with R rows per thread.
template <typename T>
static StageArgs<T> make_stage_args(const lee_handle *h, const Block &b, int stage, T *const *Y, T *rho, int R,
                                    const StageScalars &sc) {
  StageArgs<T> a{};
  a.ni = b.ni;
  a.nj = b.nj;
  a.ntx = (b.ni + ST_TX - 1) / ST_TX;
  a.nty = (b.nj + 4 * R - 1) / (4 * R);
  a.ntiles = a.ntx * a.nty;
  a.chunk = (a.ntiles + 7) / 8;
  // stage 0 reads Y0 and writes Y1 (accumulator) and Y2; odd stages read
  // Y2 and write Y3, even stages read Y3 and write Y2; y_n stays in Y0.
  // stage < 0 (lee_compute_rhs): y in Y[2], zeroed accumulator Y[3] with
  // h = b = 1 receives the RHS, Y[1] is scratch
  const int in_idx = stage == 0 ? 0 : (stage % 2 == 1 ? 2 : 3);
  const int out_idx = stage % 2 == 0 ? 2 : 3;
  a.yin = Y[in_idx];
  a.yn = Y[0];
  a.acc = Y[1];
  a.yout = Y[out_idx];
  if (stage < 0) {
    a.yin = Y[2];
    a.yn = Y[2];
    a.acc = Y[3];
    a.yout = Y[1];
  }
  a.met = b.met;
  a.jac = b.jac;
  a.iblank = b.iblank;
  a.code = b.code_fd;
  a.clean = b.clean_fd;
  a.mean = b.mean;
  a.mgrad = b.mgrad;
  a.rhob = b.rhobar;
  a.mc = sizeof(T) == 8 ? (const T *)b.mc64 : (const T *)b.mc32;
  a.rho_in = rho;
  a.rho_out = rho;
  a.u0 = b.umean[IUX];
  a.v0 = b.umean[IUY];
  a.gamma = h->cfg.gamma;
  a.axi = h->cfg.axisym;
  a.gi = 1.0 / h->cfg.gamma;
  a.gp0 = h->cfg.gamma * b.umean[IP];
  {
    const double pbar = b.umean[IP], sbar = b.umean[IS];
    a.rhob0 = std::pow(h->cfg.gamma * pbar, 1.0 / h->cfg.gamma) * std::exp(-sbar);
  }
  fill_zones(a, b);
  a.si0 = b.src.i0;
  a.si1 = b.src.i1;
  a.sj0 = b.src.j0;
  a.sj1 = b.src.j1;
  a.g = b.src.g;
  a.s1 = sc.s1;
  a.s2 = sc.s2;
  a.he = sc.he;
  a.h = (T)sc.h;
  a.b = (T)sc.b;
  a.a = (T)sc.a;
  a.fdw = h->fdw;
  return a;
}

template <typename T>
static hipError_t fused_stage_t(hipStream_t s, lee_handle *h, Block &b, int stage, const StageScalars &sc, T *const *Y,
                                T *rho) {
  const int kind = stage_kind(h, stage);
  const bool wide = h->cfg.fd_order == 11;
  const int R = wide ? 1 : b.tune_r;  // the 11-point rows run on tiles of 4 rows
  const StageArgs<T> a = make_stage_args<T>(h, b, stage, Y, rho, R, sc);
  const dim3 grid(a.chunk * 8), block(64, 4);
  const bool mf = !b.uniform_mean, ho = b.has_holes;
#define LEE_LAUNCH(MF, HO, K)                                                                     \
  do {                                                                                            \
    if (wide) hipLaunchKernelGGL((k_fused_stage<T, MF, HO, K, 1, 11>), grid, block, 0, s, a);    \
    else if (R == 1) hipLaunchKernelGGL((k_fused_stage<T, MF, HO, K, 1>), grid, block, 0, s, a); \
    else if (R == 2) hipLaunchKernelGGL((k_fused_stage<T, MF, HO, K, 2>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_fused_stage<T, MF, HO, K, 4>), grid, block, 0, s, a);             \
  } while (0)
#define LEE_KINDS(MF, HO)                                      \
  do {                                                         \
    if (kind == STAGE_FIRST) LEE_LAUNCH(MF, HO, STAGE_FIRST);  \
    else if (kind == STAGE_MID) LEE_LAUNCH(MF, HO, STAGE_MID); \
    else LEE_LAUNCH(MF, HO, STAGE_LAST);                       \
  } while (0)
  if (mf && ho) LEE_KINDS(true, true);
  else if (mf) LEE_KINDS(true, false);
  else if (ho) LEE_KINDS(false, true);
  else LEE_KINDS(false, false);
#undef LEE_KINDS
#undef LEE_LAUNCH
  if (b.carry_w)
    hipLaunchKernelGGL(k_passive_w<T>, dim3(nblocks(b.n)), dim3(TPB), 0, s, a, kind);
  if (const int nsc = h->cfg.num_scalar) {
    const dim3 sg(nblocks((size_t)nsc * b.n));
#define LEE_SCALARS(K)                                                                              \
  do {                                                                                              \
    if (wide) hipLaunchKernelGGL((k_scalar_stage<T, K, 11>), sg, dim3(TPB), 0, s, a, nsc);          \
    else hipLaunchKernelGGL((k_scalar_stage<T, K, 4>), sg, dim3(TPB), 0, s, a, nsc);                \
  } while (0)
    if (kind == STAGE_FIRST) LEE_SCALARS(STAGE_FIRST);
    else if (kind == STAGE_MID) LEE_SCALARS(STAGE_MID);
    else LEE_SCALARS(STAGE_LAST);
#undef LEE_SCALARS
  }
  return hipGetLastError();
}

hipError_t fused_stage(hipStream_t s, lee_handle *h, Block &b, int stage, const StageScalars &sc) {
  if (h->cfg.precision == LEE_FP32) return fused_stage_t<float>(s, h, b, stage, sc, b.Y32, b.rho32);
  return fused_stage_t<double>(s, h, b, stage, sc, b.Y64, b.rho64);
}

// Whether a 2-D handle's two blocks can share one stage launch and one
// filter launch: the same kernel instance (holes, mean kind, operator) and
// the same tile height; the filter also needs both blocks in the same
// stacked column-march form.
bool pair_stage_ok(const lee_handle *h) {
  if (h->nblk != 2 || h->batch || h->cfg.num_dim != 2 || h->cfg.fd_order != 4 || h->cfg.path != LEE_PATH_FUSED)
    return false;
  const Block &b0 = h->b[0], &b1 = h->b[1];
  return b0.has_holes == b1.has_holes && b0.uniform_mean == b1.uniform_mean && b0.tune_r == b1.tune_r;
}

bool pair_filter_ok(const lee_handle *h) {
  if (!pair_stage_ok(h) || h->cfg.filter_order == 11 || h->cfg.filter_order == 0) return false;
  const Block &b0 = h->b[0], &b1 = h->b[1];
  return b0.tune_mc > 1 && b0.tune_mc == b1.tune_mc && b0.tune_ml == b1.tune_ml && !b0.carry_w && !b1.carry_w;
}

template <typename T>
static hipError_t fused_stage_pair_t(hipStream_t s, lee_handle *h, int stage, const StageScalars &sc) {
  Block &b0 = h->b[0], &b1 = h->b[1];
  const int kind = stage_kind(h, stage), R = b0.tune_r;
  T *const *Y0 = sizeof(T) == 8 ? (T *const *)b0.Y64 : (T *const *)b0.Y32;
  T *const *Y1 = sizeof(T) == 8 ? (T *const *)b1.Y64 : (T *const *)b1.Y32;
  T *r0 = sizeof(T) == 8 ? (T *)b0.rho64 : (T *)b0.rho32;
  T *r1 = sizeof(T) == 8 ? (T *)b1.rho64 : (T *)b1.rho32;
  const StageArgs<T> a0 = make_stage_args<T>(h, b0, stage, Y0, r0, R, sc);
  const StageArgs<T> a1 = make_stage_args<T>(h, b1, stage, Y1, r1, R, sc);
  const dim3 grid((a0.ntiles + a1.ntiles + 7) / 8 * 8), block(64, 4);
  const bool mf = !b0.uniform_mean, ho = b0.has_holes;
#define LEE_PL(MF, HO, K)                                                                              \
  do {                                                                                                 \
    if (R == 1) hipLaunchKernelGGL((k_fused_stage_pair<T, MF, HO, K, 1>), grid, block, 0, s, a0, a1); \
    else if (R == 2) hipLaunchKernelGGL((k_fused_stage_pair<T, MF, HO, K, 2>), grid, block, 0, s, a0, a1); \
    else hipLaunchKernelGGL((k_fused_stage_pair<T, MF, HO, K, 4>), grid, block, 0, s, a0, a1);         \
  } while (0)
#define LEE_PK(MF, HO)                                     \
  do {                                                     \
    if (kind == STAGE_FIRST) LEE_PL(MF, HO, STAGE_FIRST);  \
    else if (kind == STAGE_MID) LEE_PL(MF, HO, STAGE_MID); \
    else LEE_PL(MF, HO, STAGE_LAST);                       \
  } while (0)
  if (mf && ho) LEE_PK(true, true);
  else if (mf) LEE_PK(true, false);
  else if (ho) LEE_PK(false, true);
  else LEE_PK(false, false);
#undef LEE_PK
#undef LEE_PL
  // the per-block extras of fused_stage_t: a carried w' and the passive scalars
  for (int ib = 0; ib < 2; ib++) {
    Block &b = h->b[ib];
    const StageArgs<T> &a = ib == 0 ? a0 : a1;
    if (b.carry_w) hipLaunchKernelGGL(k_passive_w<T>, dim3(nblocks(b.n)), dim3(TPB), 0, s, a, kind);
    if (const int nsc = h->cfg.num_scalar) {
      const dim3 sg(nblocks((size_t)nsc * b.n));
      if (kind == STAGE_FIRST) hipLaunchKernelGGL((k_scalar_stage<T, STAGE_FIRST, 4>), sg, dim3(TPB), 0, s, a, nsc);
      else if (kind == STAGE_MID) hipLaunchKernelGGL((k_scalar_stage<T, STAGE_MID, 4>), sg, dim3(TPB), 0, s, a, nsc);
      else hipLaunchKernelGGL((k_scalar_stage<T, STAGE_LAST, 4>), sg, dim3(TPB), 0, s, a, nsc);
    }
  }
  return hipGetLastError();
}

hipError_t fused_stage_pair(hipStream_t s, lee_handle *h, int stage, const StageScalars &sc) {
  if (h->cfg.precision == LEE_FP32) return fused_stage_pair_t<float>(s, h, stage, sc);
  return fused_stage_pair_t<double>(s, h, stage, sc);
}

// Operands of the filter on block b with tile height TYF (four variables).
template <typename T>
static FilterArgs<T> make_filter_args(const lee_handle *h, const Block &b, T *const *Y, int src, int dst, int TYF) {
  FilterArgs<T> a{};
  a.ni = b.ni;
  a.nj = b.nj;
  a.ntx = (b.ni + FL_TX - 1) / FL_TX;
  a.nty = (b.nj + TYF - 1) / TYF;
  a.ntiles = a.ntx * a.nty;
  a.chunk = (a.ntiles + 7) / 8;
  a.src = Y[src];
  a.dst = Y[dst];
  a.iblank = b.iblank;
  a.code = b.code_flt;
  a.clean = b.clean_flt;
  a.blend = (T)h->cfg.filter_blend;
  a.rem = (T)(1.0 - h->cfg.filter_blend);
  a.vars[0] = IS;
  a.vars[1] = IUX;
  a.vars[2] = IUY;
  a.vars[3] = IP;
  a.flw = h->flw;
  return a;
}

template <typename T>
static hipError_t fused_filter_t(hipStream_t s, lee_handle *h, Block &b, T *const *Y, int src, int dst) {
  const int TYF = b.tune_fty;
  FilterArgs<T> a = make_filter_args<T>(h, b, Y, src, dst, TYF);
  const dim3 grid(a.chunk * 8), block(64, 4);
  // persistent pipelined form: tune_fgrid workgroups (a multiple of 8) stride over the tiles
  const bool pipe = b.tune_fgrid > 0;
  const dim3 pgrid(std::min(a.chunk * 8, (b.tune_fgrid + 7) / 8 * 8));
  const bool o8 = h->cfg.filter_order == 8;
  if (h->cfg.filter_order == 11) {  // SFO11P: the tile form, tiles of 4 or 8 rows, halo 6
    for (int pass = 0; pass < (b.carry_w ? 2 : 1); pass++) {
      const bool four = pass == 0;
      if (!four) a.vars[0] = IUZ;
#define LEE_FLW(HO, NVV)                                                                                  \
  do {                                                                                                    \
    if (TYF == 4) hipLaunchKernelGGL((k_fused_filter<T, HO, 11, NVV, 4>), grid, block, 0, s, a);          \
    else hipLaunchKernelGGL((k_fused_filter<T, HO, 11, NVV, 8>), grid, block, 0, s, a);                   \
  } while (0)
      if (b.has_holes) {
        if (four) LEE_FLW(true, 4);
        else LEE_FLW(true, 1);
      } else {
        if (four) LEE_FLW(false, 4);
        else LEE_FLW(false, 1);
      }
#undef LEE_FLW
    }
    return hipGetLastError();
  }
  if (b.tune_mc > 0 && (!b.has_holes || b.tune_mc > 1)) {
    // column-march forms: every carried variable in one launch
    const int nv = b.carry_w ? 5 : 4;
    a.vars[4] = IUZ;
    const int NW = b.tune_mc, L = b.tune_ml;
    a.ntx = (b.ni + March::STRIDE - 1) / March::STRIDE;
    a.nty = (b.nj + NW * L - 1) / (NW * L);
    a.ntiles = nv * a.ntx * a.nty;  // waves (single-wave form) or stacks
    a.chunk = NW == 1 ? ((a.ntiles + 3) / 4 + 7) / 8 : (a.ntiles + 7) / 8;
    const dim3 mgrid(a.chunk * 8), sblock(64, NW == 1 ? 4 : NW);
#define LEE_FLM(OR, LL) hipLaunchKernelGGL((k_fused_filter_m<T, OR, LL>), mgrid, sblock, 0, s, a)
#define LEE_FLS(OR, NWV, RR)                                                                              \
  do {                                                                                                    \
    if (b.has_holes) hipLaunchKernelGGL((k_fused_filter_s<T, OR, NWV, RR, true>), mgrid, sblock, 0, s, a); \
    else hipLaunchKernelGGL((k_fused_filter_s<T, OR, NWV, RR, false>), mgrid, sblock, 0, s, a);            \
  } while (0)
#define LEE_FLM_O(OR)                              \
  do {                                             \
    if (NW == 1 && L == 8) LEE_FLM(OR, 8);         \
    else if (NW == 1) LEE_FLM(OR, 16);             \
    else if (NW == 8) LEE_FLS(OR, 8, 8);           \
    else if (L == 4) LEE_FLS(OR, 4, 4);            \
    else if (L == 8) LEE_FLS(OR, 4, 8);            \
    else LEE_FLS(OR, 4, 16);                       \
  } while (0)
    if (o8) LEE_FLM_O(8);
    else LEE_FLM_O(4);
#undef LEE_FLM_O
#undef LEE_FLS
#undef LEE_FLM
    return hipGetLastError();
  }
  // the four active variables together, then w' alone when it is carried
  for (int pass = 0; pass < (b.carry_w ? 2 : 1); pass++) {
    const bool four = pass == 0;
    if (!four) a.vars[0] = IUZ;
#define LEE_FLT1(HO, OR, NVV, TYV)                                                                       \
  do {                                                                                                   \
    if (pipe) hipLaunchKernelGGL((k_fused_filter_p<T, HO, OR, NVV, TYV>), pgrid, block, 0, s, a);        \
    else hipLaunchKernelGGL((k_fused_filter<T, HO, OR, NVV, TYV>), grid, block, 0, s, a);                \
  } while (0)
#define LEE_FLT(HO, OR, NVV)            \
  do {                                  \
    if (TYF == 4) LEE_FLT1(HO, OR, NVV, 4);  \
    else if (TYF == 8) LEE_FLT1(HO, OR, NVV, 8); \
    else if (TYF == 12) LEE_FLT1(HO, OR, NVV, 12); \
    else LEE_FLT1(HO, OR, NVV, 16);      \
  } while (0)
#define LEE_FLT_NV(HO, OR)            \
  do {                                \
    if (four) LEE_FLT(HO, OR, 4);     \
    else LEE_FLT(HO, OR, 1);          \
  } while (0)
    if (b.has_holes) {
      if (o8) LEE_FLT_NV(true, 8);
      else LEE_FLT_NV(true, 4);
    } else {
      if (o8) LEE_FLT_NV(false, 8);
      else LEE_FLT_NV(false, 4);
    }
#undef LEE_FLT_NV
#undef LEE_FLT
#undef LEE_FLT1
  }
  return hipGetLastError();
}

// The passive scalars of a LEE_SCALAR run (apply_filter filters every
// solution variable, spatial:2191-2246): one tile-form launch of one
// variable each, tiles of 8 rows (every filter form computes each point
// with the same operations).
template <typename T>
static hipError_t scalar_filter_t(hipStream_t s, lee_handle *h, Block &b, T *const *Y, int src, int dst) {
  FilterArgs<T> a = make_filter_args<T>(h, b, Y, src, dst, 8);
  const dim3 grid(a.chunk * 8), block(64, 4);
  const int o = h->cfg.filter_order;
  for (int z = 0; z < h->cfg.num_scalar; z++) {
    a.vars[0] = NV + z;
#define LEE_FLZ(HO)                                                                       \
  do {                                                                                    \
    if (o == 11) hipLaunchKernelGGL((k_fused_filter<T, HO, 11, 1, 8>), grid, block, 0, s, a); \
    else if (o == 8) hipLaunchKernelGGL((k_fused_filter<T, HO, 8, 1, 8>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_fused_filter<T, HO, 4, 1, 8>), grid, block, 0, s, a);   \
  } while (0)
    if (b.has_holes) LEE_FLZ(true);
    else LEE_FLZ(false);
#undef LEE_FLZ
  }
  return hipGetLastError();
}

// Both blocks of a pair (pair_filter_ok) in one stacked column-march launch.
template <typename T>
static hipError_t fused_filter_pair_t(hipStream_t s, lee_handle *h, int src, int dst) {
  Block &b0 = h->b[0], &b1 = h->b[1];
  const int NW = b0.tune_mc, L = b0.tune_ml;
  FilterArgs<T> a[2];
  for (int ib = 0; ib < 2; ib++) {
    Block &b = h->b[ib];
    T *const *Y = sizeof(T) == 8 ? (T *const *)b.Y64 : (T *const *)b.Y32;
    a[ib] = make_filter_args<T>(h, b, Y, src, dst, b.tune_fty);
    a[ib].ntx = (b.ni + March::STRIDE - 1) / March::STRIDE;
    a[ib].nty = (b.nj + NW * L - 1) / (NW * L);
    a[ib].ntiles = 4 * a[ib].ntx * a[ib].nty;  // stacks of the four active variables
    a[ib].chunk = (a[ib].ntiles + 7) / 8;
  }
  const dim3 grid((a[0].ntiles + a[1].ntiles + 7) / 8 * 8), block(64, NW);
  const bool ho = b0.has_holes, o8 = h->cfg.filter_order == 8;
#define LEE_FP(OR, NWV, RR)                                                                                      \
  do {                                                                                                           \
    if (ho) hipLaunchKernelGGL((k_fused_filter_s_pair<T, OR, NWV, RR, true>), grid, block, 0, s, a[0], a[1]);   \
    else hipLaunchKernelGGL((k_fused_filter_s_pair<T, OR, NWV, RR, false>), grid, block, 0, s, a[0], a[1]);     \
  } while (0)
#define LEE_FP_O(OR)                    \
  do {                                  \
    if (NW == 8) LEE_FP(OR, 8, 8);      \
    else if (L == 4) LEE_FP(OR, 4, 4);  \
    else if (L == 8) LEE_FP(OR, 4, 8);  \
    else LEE_FP(OR, 4, 16);             \
  } while (0)
  if (o8) LEE_FP_O(8);
  else LEE_FP_O(4);
#undef LEE_FP_O
#undef LEE_FP
  hipError_t e = hipGetLastError();
  for (int ib = 0; ib < 2 && e == hipSuccess && h->cfg.num_scalar; ib++) {
    Block &b = h->b[ib];
    T *const *Y = sizeof(T) == 8 ? (T *const *)b.Y64 : (T *const *)b.Y32;
    e = scalar_filter_t<T>(s, h, b, Y, src, dst);
  }
  return e;
}

hipError_t fused_filter_pair(hipStream_t s, lee_handle *h, int src, int dst) {
  if (h->cfg.precision == LEE_FP32) return fused_filter_pair_t<float>(s, h, src, dst);
  return fused_filter_pair_t<double>(s, h, src, dst);
}

hipError_t fused_filter(hipStream_t s, lee_handle *h, Block &b, int src, int dst) {
  const bool f32 = h->cfg.precision == LEE_FP32;
  hipError_t e = f32 ? fused_filter_t<float>(s, h, b, b.Y32, src, dst) : fused_filter_t<double>(s, h, b, b.Y64, src, dst);
  if (e != hipSuccess || h->cfg.num_scalar == 0) return e;
  return f32 ? scalar_filter_t<float>(s, h, b, b.Y32, src, dst) : scalar_filter_t<double>(s, h, b, b.Y64, src, dst);
}

// ------------------------------------------------------------ batched ensemble

// A device table as a constant-address-space pointer (the kernels' view).
template <typename P>
static const LEE_CONST_AS P *CA(const P *p) {
  return (const LEE_CONST_AS P *)p;
}

void batch_release(lee_handle *h) {
  BatchPlan &p = h->bp;
  if (p.exec) (void)hipGraphExecDestroy(p.exec);
  p.exec = nullptr;
  for (void *q : {p.stage_args, p.filter_args, (void *)p.tile_map_stage, (void *)p.tile_map_filter})
    if (q) (void)hipFree(q);
  p.stage_args = p.filter_args = nullptr;
  p.tile_map_stage = p.tile_map_filter = nullptr;
  p.ready = false;
}

// Uploads the per-solve operand tables of every stage and of the filter for
// the current tile shape, and the tile prefix sums.  The schedule and the
// step counter are allocated by the caller.
// With p.tuning the tables describe runs on scratch buffers only (every
// stage: the right-hand-side form, Y2 -> Y3 and Y1; filter: Y1 -> Y2), so
// the autotuner can time launches without touching the state in Y0.
template <typename T>
static void batch_resident_t(lee_handle *h);

template <typename T>
static hipError_t batch_prepare_t(hipStream_t s, lee_handle *h) {
  BatchPlan &p = h->bp;
  const int ns = h->nblk, S = h->cfg.rk_stages;
  std::vector<StageArgs<T>> sa((size_t)S * ns);
  std::vector<FilterArgs<T>> fa(ns);
  std::vector<int> ts, tf;
  const StageScalars zero{0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < ns; k++) {
    Block &b = h->b[k];
    T *const *Y = sizeof(T) == 8 ? (T *const *)b.Y64 : (T *const *)b.Y32;
    T *rho = sizeof(T) == 8 ? (T *)b.rho64 : (T *)b.rho32;
    for (int st = 0; st < S; st++)
      sa[(size_t)st * ns + k] = make_stage_args<T>(h, b, p.tuning ? -1 : st, Y, rho, p.R, zero);
    fa[k] = make_filter_args<T>(h, b, Y, 1, p.tuning ? 2 : 0, p.TY);
    if (p.stack) {  // stacks of four waves of p.stack rows, four variables
      fa[k].ntx = (b.ni + March::STRIDE - 1) / March::STRIDE;
      fa[k].nty = (b.nj + 4 * p.stack - 1) / (4 * p.stack);
      fa[k].ntiles = 4 * fa[k].ntx * fa[k].nty;
    }
    for (int t = 0; t < sa[k].ntiles; t++) {
      ts.push_back(k);
      ts.push_back(t);
    }
    for (int t = 0; t < fa[k].ntiles; t++) {
      tf.push_back(k);
      tf.push_back(t);
    }
  }
  p.total_stage = (int)(ts.size() / 2);
  p.total_filter = (int)(tf.size() / 2);
  hipError_t e;
  if ((e = hipMalloc(&p.stage_args, sizeof(StageArgs<T>) * sa.size())) != hipSuccess) return e;
  if ((e = hipMalloc(&p.filter_args, sizeof(FilterArgs<T>) * fa.size())) != hipSuccess) return e;
  if ((e = hipMalloc((void **)&p.tile_map_stage, sizeof(int) * ts.size())) != hipSuccess) return e;
  if ((e = hipMalloc((void **)&p.tile_map_filter, sizeof(int) * tf.size())) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(p.stage_args, sa.data(), sizeof(StageArgs<T>) * sa.size(), hipMemcpyHostToDevice, s)))
    return e;
  if ((e = hipMemcpyAsync(p.filter_args, fa.data(), sizeof(FilterArgs<T>) * fa.size(), hipMemcpyHostToDevice, s)))
    return e;
  if ((e = hipMemcpyAsync(p.tile_map_stage, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice, s)))
    return e;
  if ((e = hipMemcpyAsync(p.tile_map_filter, tf.data(), sizeof(int) * tf.size(), hipMemcpyHostToDevice, s)))
    return e;
  // the host vectors die here: wait for the copies
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  batch_resident_t<T>(h);
  p.ready = true;
  return hipSuccess;
}

hipError_t batch_prepare(hipStream_t s, lee_handle *h) {
  batch_release(h);
  if (h->cfg.precision == LEE_FP32) return batch_prepare_t<float>(s, h);
  return batch_prepare_t<double>(s, h);
}

// Grid of a batch launch over `total` list entries: one workgroup per entry
// (setting 0, or a setting at least that large), a persistent grid of the
// given size rounded up to a multiple of 8, or (LEE_GRID_RESIDENT) the
// workgroups the device holds at once for that kernel.
static unsigned batch_grid(int total, int setting, int resident) {
  const int chunk = (total + 7) / 8;
  const int full = 8 * chunk;
  const int g = setting == LEE_GRID_RESIDENT ? resident : setting;
  if (g <= 0 || g >= full) return (unsigned)full;
  return (unsigned)((g + 7) / 8 * 8);
}

template <typename T>
using BatchStageFn = void (*)(BatchStageArgs<T>, const LEE_CONST_AS StageArgs<T> *, const LEE_CONST_AS int *,
                              const LEE_CONST_AS BatchSched *, const int *, int *);
template <typename T>
using BatchFilterFn = void (*)(BatchFilterArgs<T>, const LEE_CONST_AS FilterArgs<T> *, const LEE_CONST_AS int *);

// queued: the persistent work-queue kernel (a grid smaller than the list)
template <typename T, bool MF, int K>
static BatchStageFn<T> batch_stage_fn_r(int R, bool queued) {
  if (queued) {
    if (R == 1) return k_batch_stage_q<T, MF, K, 1>;
    if (R == 2) return k_batch_stage_q<T, MF, K, 2>;
    return k_batch_stage_q<T, MF, K, 4>;
  }
  if (R == 1) return k_batch_stage<T, MF, K, 1>;
  if (R == 2) return k_batch_stage<T, MF, K, 2>;
  return k_batch_stage<T, MF, K, 4>;
}

template <typename T, bool MF>
static BatchStageFn<T> batch_stage_fn_k(int kind, int R, bool queued) {
  if (kind == STAGE_FIRST) return batch_stage_fn_r<T, MF, STAGE_FIRST>(R, queued);
  if (kind == STAGE_MID) return batch_stage_fn_r<T, MF, STAGE_MID>(R, queued);
  return batch_stage_fn_r<T, MF, STAGE_LAST>(R, queued);
}

template <typename T>
static BatchStageFn<T> batch_stage_fn(const lee_handle *h, int kind, bool queued) {
  const int R = h->bp.R;
  return h->b[0].uniform_mean ? batch_stage_fn_k<T, false>(kind, R, queued)
                              : batch_stage_fn_k<T, true>(kind, R, queued);
}

template <typename T, int OR>
static BatchFilterFn<T> batch_filter_fn_o(const BatchPlan &p) {
  if (p.stack == 4) return k_batch_filter_s<T, OR, 4>;
  if (p.stack == 8) return k_batch_filter_s<T, OR, 8>;
  if (p.stack) return k_batch_filter_s<T, OR, 16>;
  if (p.TY == 4) return k_batch_filter<T, OR, 4>;
  if (p.TY == 8) return k_batch_filter<T, OR, 8>;
  if (p.TY == 12) return k_batch_filter<T, OR, 12>;
  return k_batch_filter<T, OR, 16>;
}

template <typename T>
static BatchFilterFn<T> batch_filter_fn(const lee_handle *h) {
  return h->cfg.filter_order == 8 ? batch_filter_fn_o<T, 8>(h->bp) : batch_filter_fn_o<T, 4>(h->bp);
}

// Workgroups of kernel f resident on the whole device at once (its
// occupancy per CU, from registers and LDS, times the CUs), a multiple of 8.
static int resident_workgroups(const void *f, int device) {
  int per = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, TPB, 0) != hipSuccess || per < 1) per = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 8;
  return std::max(8, per * cus / 8 * 8);
}

template <typename T>
static void batch_resident_t(lee_handle *h) {
  BatchPlan &p = h->bp;
  int rs = 1 << 30;
  for (int kind : {STAGE_FIRST, STAGE_MID, STAGE_LAST})
    rs = std::min(rs, resident_workgroups(reinterpret_cast<const void *>(batch_stage_fn<T>(h, kind, true)), h->device));
  p.res_stage = rs;
  p.res_filter = resident_workgroups(reinterpret_cast<const void *>(batch_filter_fn<T>(h)), h->device);
}

template <typename T>
static hipError_t batch_stage_t(hipStream_t s, lee_handle *h, int stage) {
  BatchPlan &p = h->bp;
  const int ns = h->nblk;
  BatchStageArgs<T> ba{};
  ba.args = (const StageArgs<T> *)p.stage_args + (size_t)stage * ns;
  ba.map = p.tile_map_stage;
  ba.sched = p.sched;
  ba.counter = p.counter;
  ba.nsolve = ns;
  ba.nstage = h->cfg.rk_stages;
  ba.stage = stage;
  ba.total = p.total_stage;
  ba.chunk = (p.total_stage + 7) / 8;
  ba.b = (T)h->b_i[stage];
  ba.a = (T)h->a_next[stage];
  const int kind = p.tuning ? STAGE_MID : stage_kind(h, stage);
  const unsigned g = batch_grid(p.total_stage, p.grid_stage, p.res_stage);
  const bool queued = g < 8u * (unsigned)ba.chunk;
  const dim3 grid(g), block(64, 4);
  hipLaunchKernelGGL(batch_stage_fn<T>(h, kind, queued), grid, block, 0, s, ba, CA(ba.args), CA(ba.map), CA(ba.sched),
                     ba.counter, p.counter + QSTRIDE);
  return hipGetLastError();
}

hipError_t batch_stage(hipStream_t s, lee_handle *h, int stage) {
  if (h->cfg.precision == LEE_FP32) return batch_stage_t<float>(s, h, stage);
  return batch_stage_t<double>(s, h, stage);
}

template <typename T>
static hipError_t batch_filter_t(hipStream_t s, lee_handle *h) {
  BatchPlan &p = h->bp;
  BatchFilterArgs<T> ba{};
  ba.args = (const FilterArgs<T> *)p.filter_args;
  ba.map = p.tile_map_filter;
  ba.counter = p.tuning ? p.counter + 1 : p.counter;  // tuning runs leave the step counter alone
  ba.nsolve = h->nblk;
  ba.total = p.total_filter;
  ba.chunk = (p.total_filter + 7) / 8;
  const dim3 grid(batch_grid(p.total_filter, p.grid_filter, p.res_filter)), block(64, 4);
  hipLaunchKernelGGL(batch_filter_fn<T>(h), grid, block, 0, s, ba, CA(ba.args), CA(ba.map));
  return hipGetLastError();
}

hipError_t batch_filter(hipStream_t s, lee_handle *h) {
  if (h->cfg.precision == LEE_FP32) return batch_filter_t<float>(s, h);
  return batch_filter_t<double>(s, h);
}

hipError_t cast_state(hipStream_t s, const Block &, int, const double *src, float *dst, size_t n) {
  hipLaunchKernelGGL((k_cast<double, float>), dim3(nblocks(n)), dim3(TPB), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t cast_back(hipStream_t s, const Block &, int, const float *src, double *dst, size_t n) {
  hipLaunchKernelGGL((k_cast<float, double>), dim3(nblocks(n)), dim3(TPB), 0, s, src, dst, n);
  return hipGetLastError();
}

}  // namespace lee
