// Discrete adjoint of the fused 2-D LEE time advance (no reference
// counterpart: the reference has no adjoint).
//
// The forward step with stage inputs u_i, stage RHS k_i = L(u_i, rho_i) + q s_i,
// accumulator acc_{i+1} = acc_i + b_i h k_i (acc_0 = y_n), next input
// u_{i+1} = y_n + a_i h k_i, filtered state y_{n+1} = F(acc_S) and the
// stale aux density rho_{n+1} = R(acc_S) (R(y) = rho_bar (p'/(gamma p_bar) - s'),
// consumed by stage 0 of the next step) is linear in (y_n, rho_n) and
// affine in the source field q.  Its exact transpose, step by step
// backwards, is
//   lambda* = F^T lambda_{n+1} + R^T mu_{n+1}
//   kappa_i = b_i h lambda* + a_i h lambda_u{i+1}        (last stage: b_i h lambda*)
//   lambda_u_i = L_u^T kappa_i + R^T L_rho^T kappa_i     (i >= 1)
//   lambda_n = lambda* + sum_{i>=1} lambda_u_i + L_u^T kappa_0,   mu_n = L_rho^T kappa_0
//   dJ/dq += sum_i s_i kappa_i[p']
// with the transposed operators built from the same row codes as the
// forward: (D^T f)_i = sum over rows j touching i of w(code_j, i - j) f_j.
// L^T applies D_xi^T and D_eta^T to per-point flux fields
// J (m_xi,x A + m_xi,y B), J (m_eta,x A + m_eta,y B), where A, B are the
// coefficients of d/dx and d/dy in each RHS term times kappa, evaluated at
// the row's own point.  The filter transposes as rem I + blend F_xi^T F_eta^T.
#include <hip/hip_runtime.h>

#include <cmath>

#include "lee_device.h"
#include "lee_internal.h"

namespace lee {

namespace {

// Weight of the FD row with code `code` (at point j) on point j + d: the
// rows of fd_row (lee_device.h) written out by offset.
template <typename T>
__device__ __forceinline__ T fd_wt(int code, int d) {
  const T c1 = (T)(8.0 / 12.0), c2 = (T)(-1.0 / 12.0);
  switch (code) {
    case CODE_INTERIOR:
      return d == -2 ? -c2 : d == -1 ? -c1 : d == 1 ? c1 : d == 2 ? c2 : (T)0;
    case CODE_ZERO:
      return (T)0;
    case CODE_LEFT + 0:
      return d == 0 ? (T)(-3.0 / 2.0) : d == 1 ? (T)(4.0 / 2.0) : d == 2 ? (T)(-1.0 / 2.0) : (T)0;
    case CODE_LEFT + 1:
      return d == -1 ? (T)(-1.0 / 2.0) : d == 1 ? (T)(1.0 / 2.0) : (T)0;
    case CODE_LEFT + 2:
      return d == 0 ? (T)(3.0 / 2.0) : d == -1 ? (T)(-4.0 / 2.0) : d == -2 ? (T)(1.0 / 2.0) : (T)0;
    default:
      return d == 1 ? (T)(1.0 / 2.0) : d == -1 ? (T)(-1.0 / 2.0) : (T)0;
  }
}

// Transposed FD row at coordinate c of a line of length len (no holes):
// the rows of the neighbours c + o, o in [-2, 2], that exist.
template <typename T>
__device__ __forceinline__ T fd_row_t(const T *f, int st, int c, int len, bool inner) {
  if (inner) {  // every neighbour row is interior: D^T = -D there
    const T c1 = (T)(8.0 / 12.0), c2 = (T)(-1.0 / 12.0);
    T acc = c2 * f[-2 * st];
    acc += c1 * f[-st];
    acc += -c1 * f[st];
    acc += -c2 * f[2 * st];
    return acc;
  }
  T acc = 0;
#pragma unroll
  for (int o = -2; o <= 2; o++) {
    const int j = c + o;
    if (j < 0 || j >= len) continue;
    acc += fd_wt<T>(default_code(j, len, 2), -o) * f[o * st];
  }
  return acc;
}

// Transposed FD row at block point p (coordinate c of a line of length
// len, stride gs in the block, st in LDS) with the per-point codes of a
// block with holes: the rows of the neighbours c + o, o in [-2, 2].
template <typename T>
__device__ __forceinline__ T fd_row_t_codes(const T *f, int st, int c, int len, const int8_t *code, int p, int gs) {
  T acc = 0;
#pragma unroll
  for (int o = -2; o <= 2; o++) {
    const int j = c + o;
    if (j < 0 || j >= len) continue;
    acc += fd_wt<T>((int)code[p + o * gs], -o) * f[o * st];
  }
  return acc;
}

// Weight of the filter row with code `code` on offset d (rows are centred
// and symmetric; filter_row in lee_device.h).
template <typename T, int ORDER>
__device__ __forceinline__ T flt_wt(int code, int d) {
  const int ad = d < 0 ? -d : d;
  if (ORDER == 8) {
    if (code == CODE_INTERIOR) {
      const T c[5] = {(T)(93.0 / 128.0), (T)(7.0 / 32.0), (T)(-7.0 / 64.0), (T)(1.0 / 32.0), (T)(-1.0 / 256.0)};
      return ad <= 4 ? c[ad] : (T)0;
    }
    if (code == CODE_ZERO) return (T)0;
    const int k = code - CODE_LEFT;
    const int row = k < 4 ? k : k - 4;
    if (row <= 1) return ad == 0 ? (T)1 : (T)0;
    if (row == 2) {
      const T c[3] = {(T)(5.0 / 8.0), (T)((1.0 / 2.0) / 2.0), (T)((-1.0 / 8.0) / 2.0)};
      return ad <= 2 ? c[ad] : (T)0;
    }
    const T c[4] = {(T)(11.0 / 16.0), (T)((15.0 / 32.0) / 2.0), (T)((-3.0 / 16.0) / 2.0), (T)((1.0 / 32.0) / 2.0)};
    return ad <= 3 ? c[ad] : (T)0;
  } else {
    if (code == CODE_INTERIOR) {
      const T c[3] = {(T)(5.0 / 8.0), (T)((1.0 / 2.0) / 2.0), (T)((-1.0 / 8.0) / 2.0)};
      return ad <= 2 ? c[ad] : (T)0;
    }
    if (code == CODE_ZERO) return (T)0;
    return ad == 0 ? (T)1 : (T)0;
  }
}

template <typename T, int ORDER>
__device__ __forceinline__ T flt_row_t(const T *f, int st, int c, int len, bool inner) {
  constexpr int NB = ORDER / 2, H = ORDER / 2;
  if (inner) {  // symmetric interior rows: F^T = F
    return filter_row<T, ORDER>(f, st, CODE_INTERIOR);
  }
  T acc = 0;
#pragma unroll
  for (int o = -H; o <= H; o++) {
    const int j = c + o;
    if (j < 0 || j >= len) continue;
    acc += flt_wt<T, ORDER>(default_code(j, len, NB), -o) * f[o * st];
  }
  return acc;
}

template <typename T, int ORDER>
__device__ __forceinline__ T flt_row_t_codes(const T *f, int st, int c, int len, const int8_t *code, int p, int gs) {
  constexpr int H = ORDER / 2;
  T acc = 0;
#pragma unroll
  for (int o = -H; o <= H; o++) {
    const int j = c + o;
    if (j < 0 || j >= len) continue;
    acc += flt_wt<T, ORDER>((int)code[p + o * gs], -o) * f[o * st];
  }
  return acc;
}

template <typename T>
struct AdjArgs {
  int ni, nj, ntx, nty, ntiles, chunk;
  const T *lacc;   // adjoint of the accumulator after the last stage (5n)
  const T *lin;    // adjoint of the next stage input (null at the last stage)
  const T *lyin;   // running adjoint of y_n (null when nothing has reached it yet)
  T *lout;         // adjoint of this stage's input; at stage 0 the adjoint of y_n
  T *lyout;        // running adjoint of y_n (middle stages; may equal lyin)
  T *mu;           // adjoint of the stale aux density (stage 0, non-uniform mean)
  double *grad;    // adjoint of the source field (n), accumulated (null: none)
  const double *met, *jac, *mean, *mgrad, *rhob;
  double u0, v0, gp0, rhob0, gamma, gi;
  int nz;
  ZoneT<T> z[MAX_ZONES];
  double hb, ha;   // b_i h and a_i h
  double src;      // s1 s2 of this stage
  const int8_t *iblank, *code;  // holes only: iblank, FD codes (xi then eta)
  const int *clean;             // holes only: integral image of points off the interior FD rows
  const T *rad;                 // axisymmetric runs: the forward's radius field r + 1e-99 (null: none)
};

template <typename T, int R>
using AdjLds = T[8][4 * R + 2 * ST_H][ST_TX + 2 * ST_H];

// One reversed RK stage on a 64 x 4R tile: kappa and the flux fields on the
// tile and its 2-point halo go to LDS, then the transposed stencils, the
// point-local transposed terms and the accumulations of the running
// adjoints for the R points of each thread.
// Blocks with holes (HOLES): the forward stage writes 0 at BLANKED points
// whatever its inputs, so the adjoints arriving there (lacc, lin) count as 0
// and kappa vanishes there; the transposed rows take the per-point codes.
template <typename T, bool MEANF, int KIND, int R, bool HOLES>
__global__ void __launch_bounds__(TPB) k_adj_stage(AdjArgs<T> a) {
  constexpr int TY = 4 * R, NR = TY + 2 * ST_H, RS = ST_TX + 2 * ST_H;
  __shared__ AdjLds<T, R> F;  // F[v] xi-flux, F[4 + v] eta-flux
  const int tile = xcd_tile(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * ST_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const int VOFF[4] = {IS * n, IUX * n, IUY * n, IP * n};
  const bool inner = HOLES ? box_clean(a.clean, ni, nj, i0 - ST_H, i0 + ST_TX + ST_H, j0 - ST_H, j0 + TY + ST_H)
                           : (i0 >= 2 * ST_H && i0 + ST_TX + 2 * ST_H <= ni && j0 >= 2 * ST_H &&
                              j0 + TY + 2 * ST_H <= nj);
  const T hb = (T)a.hb, ha = (T)a.ha;

  // (1) kappa and fluxes over the tile plus halo; row r = ty + 4k
  // (k = 0..R), columns tx and, for tx < 4, 64 + tx
#pragma unroll
  for (int k = 0; k <= R; k++) {
    const int r = ty + 4 * k;
    const int gj = j0 - ST_H + r;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int c = half == 0 ? tx : ST_TX + tx;
      if (half == 1 && tx >= 2 * ST_H) continue;
      const int gi = i0 - ST_H + c;
      const bool ok = gj >= 0 && gj < nj && gi >= 0 && gi < ni && !(HOLES && a.iblank[gj * ni + gi] == BLANKED);
      double fx[4] = {0.0, 0.0, 0.0, 0.0}, fy[4] = {0.0, 0.0, 0.0, 0.0};
      if (ok) {
        const int p = gj * ni + gi;
        T kap[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
          T kv = hb * a.lacc[VOFF[v] + p];
          if (KIND != STAGE_LAST && a.lin) kv += ha * a.lin[VOFF[v] + p];
          kap[v] = kv;
        }
        double u0 = a.u0, v0 = a.v0, gp = a.gp0, irb = 1.0 / a.rhob0;
        if (MEANF) {
          u0 = a.mean[IUX * n + p];
          v0 = a.mean[IUY * n + p];
          gp = a.gamma * a.mean[IP * n + p];
          irb = 1.0 / a.rhob[p];
        }
        const double ks = kap[0], ku = kap[1], kv = kap[2], kp = kap[3];
        // coefficients of d/dx (A) and d/dy (B) of each variable
        const double A[4] = {-u0 * ks, -u0 * ku - gp * kp, -u0 * kv, -u0 * kp - ku * irb};
        const double B[4] = {-v0 * ks, -v0 * ku, -v0 * kv - gp * kp, -v0 * kp - kv * irb};
        const double mxx = a.met[p], mxy = a.met[(size_t)n + p], myx = a.met[(size_t)2 * n + p],
                     myy = a.met[(size_t)3 * n + p];
        const double J = a.jac[p];
#pragma unroll
        for (int v = 0; v < 4; v++) {
          fx[v] = (mxx * A[v] + mxy * B[v]) * J;
          fy[v] = (myx * A[v] + myy * B[v]) * J;
        }
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        F[v][r][c] = (T)fx[v];
        F[4 + v][r][c] = (T)fy[v];
      }
    }
  }
  __syncthreads();

  // (2) the R points of the thread
  const int gi = i0 + tx;
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int rr = ty + 4 * k;
    const int gj = j0 + rr;
    if (!inner && (gi >= ni || gj >= nj)) continue;
    const int p = gj * ni + gi;
    const int lj = rr + ST_H, li = tx + ST_H;
    T lam[4];
    if (HOLES && !inner) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        lam[v] = fd_row_t_codes<T>(&F[v][lj][li], 1, gi, ni, a.code, p, 1);
        lam[v] += fd_row_t_codes<T>(&F[4 + v][lj][li], RS, gj, nj, a.code + n, p, ni);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        lam[v] = fd_row_t<T>(&F[v][lj][li], 1, gi, ni, inner);
        lam[v] += fd_row_t<T>(&F[4 + v][lj][li], RS, gj, nj, inner);
      }
    }
    const bool hole = HOLES && a.iblank[p] == BLANKED;
    // kappa at the point (recomputed from the same operands)
    T kap[4], lacc[4], lin[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      lacc[v] = hole ? (T)0 : a.lacc[VOFF[v] + p];
      lin[v] = (KIND != STAGE_LAST && a.lin && !hole) ? a.lin[VOFF[v] + p] : (T)0;
      T kv = hb * lacc[v];
      if (KIND != STAGE_LAST && a.lin) kv += ha * lin[v];
      kap[v] = kv;
    }
    // sponges: r_v -= sigma y_v
#pragma unroll 1
    for (int z = 0; z < a.nz; z++) {
      const ZoneT<T> &zz = a.z[z];
      if (gi >= zz.i0 && gi <= zz.i1 && gj >= zz.j0 && gj <= zz.j1) {
        const T sg = zz.sig[(gj - zz.j0) * zz.nzi + (gi - zz.i0)];
#pragma unroll
        for (int v = 0; v < 4; v++) lam[v] += -sg * kap[v];
      }
    }
    T mur = 0;
    if (MEANF) {
      const T u0 = (T)a.mean[IUX * n + p], v0 = (T)a.mean[IUY * n + p];
      const T rb = (T)a.rhob[p];
      const T g = (T)a.gamma;
      const T gsx = (T)a.mgrad[(2 * IS + 0) * n + p], gsy = (T)a.mgrad[(2 * IS + 1) * n + p];
      const T gux = (T)a.mgrad[(2 * IUX + 0) * n + p], guy = (T)a.mgrad[(2 * IUX + 1) * n + p];
      const T gvx = (T)a.mgrad[(2 * IUY + 0) * n + p], gvy = (T)a.mgrad[(2 * IUY + 1) * n + p];
      const T gpx = (T)a.mgrad[(2 * IP + 0) * n + p], gpy = (T)a.mgrad[(2 * IP + 1) * n + p];
      const T ks = kap[0], ku = kap[1], kv = kap[2], kp = kap[3];
      lam[1] += -(gsx * ks) - gux * ku - gvx * kv - gpx * kp;
      lam[2] += -(gsy * ks) - guy * ku - gvy * kv - gpy * kp;
      lam[3] += -(g * gux * kp) - g * gvy * kp;
      mur = -((u0 * gux + v0 * guy) * ku + (u0 * gvx + v0 * gvy) * kv) / rb;
      if (a.rad) {
        // axisymmetric term r_p -= (gamma pbar v' + gamma vbar p') / r
        // (spatial:2119-2126), with the forward's gamma pbar and r fields
        const T rr = a.rad[p], gpb = (T)(a.gamma * a.mean[IP * n + p]);
        lam[2] += -(gpb / rr) * kp;
        lam[3] += -(g * v0 / rr) * kp;
      }
      if (KIND != STAGE_FIRST) {  // rho' of this stage's input: R(u)
        const double pb = a.mean[IP * n + p];
        lam[0] += -(rb * mur);
        lam[3] += (T)((double)mur * a.rhob[p] * a.gi / pb);
      }
    }
    if (a.grad) a.grad[p] += a.src * (double)kap[3];
    if (KIND == STAGE_FIRST) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        T out = lam[v] + lacc[v] + lin[v];
        if (a.lyin) out += a.lyin[VOFF[v] + p];
        a.lout[VOFF[v] + p] = out;
      }
      if (MEANF) a.mu[p] = mur;
    } else {
#pragma unroll
      for (int v = 0; v < 4; v++) a.lout[VOFF[v] + p] = lam[v];
      if (a.lyout) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          T y = lin[v];
          if (a.lyin) y += a.lyin[VOFF[v] + p];
          a.lyout[VOFF[v] + p] = y;
        }
      }
    }
  }
}

template <typename T>
struct AdjFilterArgs {
  int ni, nj, ntx, nty, ntiles, chunk;
  const T *src;    // adjoint of the filtered state (5n)
  T *dst;          // adjoint of the unfiltered state (5n)
  const T *mu;     // adjoint of the stale aux density taken from the unfiltered state (null: none)
  const double *mean, *rhob;
  double gi;
  T rem, blend;
  const int8_t *iblank, *code;  // holes only: iblank, filter codes (xi then eta)
  const int *clean;             // holes only: integral image of points off the interior filter rows
};

// Transposed filter on a 64 x TY tile: eta^T sweep of the (TY + 8) x 72
// region into LDS, then the xi^T sweep, blend, and R^T mu.
// Blocks with holes (HOLES): the forward filter is followed by the hole
// reset, so the adjoint arriving at a BLANKED point counts as 0, and the
// transposed rows take the per-point filter codes.
template <typename T, int ORDER, int TY, bool HOLES>
__global__ void __launch_bounds__(TPB) k_adj_filter(AdjFilterArgs<T> a) {
  constexpr int NR = TY + 2 * FL_H, NC = FL_TX + 2 * FL_H, KR = NR / 4;
  __shared__ T in[4][NR][NC];
  __shared__ T mid[4][TY][NC];
  const int tile = xcd_tile(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const int VOFF[4] = {IS * n, IUX * n, IUY * n, IP * n};
  // (no filter codes when the filter is off: blend is 0 and only the
  // masked identity part counts)
  const bool inner = HOLES ? (a.clean && box_clean(a.clean, ni, nj, i0 - FL_H, i0 + FL_TX + FL_H, j0 - FL_H,
                                                   j0 + TY + FL_H))
                           : (i0 >= 2 * FL_H && i0 + FL_TX + 2 * FL_H <= ni && j0 >= 2 * FL_H &&
                              j0 + TY + 2 * FL_H <= nj);
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int r = ty + 4 * k;
    const int gj = j0 - FL_H + r;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int c = half == 0 ? tx : FL_TX + tx;
      if (half == 1 && tx >= 2 * FL_H) continue;
      const int gi = i0 - FL_H + c;
      const bool ok = gj >= 0 && gj < nj && gi >= 0 && gi < ni && !(HOLES && a.iblank[gj * ni + gi] == BLANKED);
#pragma unroll
      for (int v = 0; v < 4; v++) in[v][r][c] = ok ? a.src[VOFF[v] + gj * ni + gi] : (T)0;
    }
  }
  __syncthreads();
  // eta^T over the tile rows, every column of the region
#pragma unroll
  for (int k = 0; k < TY / 4; k++) {
    const int r = ty + 4 * k;
    const int gj = j0 + r;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int c = half == 0 ? tx : FL_TX + tx;
      if (half == 1 && tx >= 2 * FL_H) continue;
      const int gi = i0 - FL_H + c;
      const bool ok = gj < nj && gi >= 0 && gi < ni;
      if (HOLES && !inner) {
#pragma unroll
        for (int v = 0; v < 4; v++)
          mid[v][r][c] = ok && a.code
                             ? flt_row_t_codes<T, ORDER>(&in[v][r + FL_H][c], NC, gj, nj, a.code + n, gj * ni + gi, ni)
                             : (T)0;
      } else {
#pragma unroll
        for (int v = 0; v < 4; v++)
          mid[v][r][c] = ok ? flt_row_t<T, ORDER>(&in[v][r + FL_H][c], NC, gj, nj, inner) : (T)0;
      }
    }
  }
  __syncthreads();
  const int gi = i0 + tx;
#pragma unroll
  for (int k = 0; k < TY / 4; k++) {
    const int r = ty + 4 * k;
    const int gj = j0 + r;
    if (gi >= ni || gj >= nj) continue;
    const int p = gj * ni + gi;
    T out[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const T f = (HOLES && !inner) ? (a.code ? flt_row_t_codes<T, ORDER>(&mid[v][r][tx + FL_H], 1, gi, ni, a.code, p, 1)
                                              : (T)0)
                                    : flt_row_t<T, ORDER>(&mid[v][r][tx + FL_H], 1, gi, ni, inner);
      out[v] = a.rem * in[v][r + FL_H][tx + FL_H] + a.blend * f;
    }
    if (a.mu) {
      const double m = (double)a.mu[p] * a.rhob[p];
      out[0] += (T)(-m);
      out[3] += (T)(m * a.gi / a.mean[IP * n + p]);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) a.dst[VOFF[v] + p] = out[v];
  }
}

// Observer functional: J_k = sum_l w_l p'(l)^2 for one state, one workgroup,
// fixed summation order.
template <typename T>
__global__ void __launch_bounds__(1024) k_obs_energy(const T *__restrict__ y, const int *__restrict__ l,
                                                     const double *__restrict__ w, int nobs, int n,
                                                     double *__restrict__ out) {
  __shared__ double part[1024];
  double acc = 0.0;
  for (int q = threadIdx.x; q < nobs; q += 1024) {
    const double pv = (double)y[(size_t)IP * n + l[q]];
    acc += w[q] * pv * pv;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += part[0];
}

// p' at the observer points of one state.
template <typename T>
__global__ void __launch_bounds__(TPB) k_obs_gather(const T *__restrict__ y, const int *__restrict__ l, int nobs,
                                                    int n, double *__restrict__ out) {
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q < nobs) out[q] = (double)y[(size_t)IP * n + l[q]];
}

// lambda_p += scale 2 w p' at the observer points.
template <typename T>
__global__ void __launch_bounds__(TPB) k_obs_seed(T *__restrict__ lam, const int *__restrict__ l,
                                                  const double *__restrict__ w, const double *__restrict__ samp,
                                                  int nobs, int n, double scale) {
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q < nobs) {
    const size_t p = (size_t)IP * n + l[q];
    lam[p] = (T)((double)lam[p] + scale * 2.0 * w[q] * samp[q]);
  }
}

// lambda += R^T mu: the aux density of the initial state, R(y) =
// rho_bar (p' gi / p_bar - s'), is the stale density of the first step.
template <typename T>
__global__ void __launch_bounds__(TPB) k_add_rt_mu(T *__restrict__ lam, const T *__restrict__ mu,
                                                   const double *__restrict__ mean, const double *__restrict__ rhob,
                                                   double gi, int n) {
  const int p = blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  const double m = (double)mu[p] * rhob[p];
  lam[(size_t)IS * n + p] = (T)((double)lam[(size_t)IS * n + p] - m);
  lam[(size_t)IP * n + p] = (T)((double)lam[(size_t)IP * n + p] + m * gi / mean[(size_t)IP * n + p]);
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_fill_t(T *__restrict__ a, size_t n) {
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p < n) a[p] = (T)0;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// ------------------------------------------- transposed boundary update
//
// The forward boundary update after a stage (spatial::update_boundary,
// spatial:2162-2187) is linear: each BC face replaces the values at its
// points, then the overset interpolation replaces every receiver by the
// weighted sum of its donors, all donor sums formed first.  Its transpose
// runs backwards: the overset part, then the faces of each block in reverse
// order.

template <typename T>
struct AdjOvTables {
  T *lam[MAX_OV_BLOCKS];
  int ni[MAX_OV_BLOCKS], nn[MAX_OV_BLOCKS];
};

// Overset transpose, part 1: each receiver's adjoint is kept in buf and
// zeroed (the forward overwrote the receiver's value).
template <typename T>
__global__ void __launch_bounds__(TPB) k_adj_ov_take(AdjOvTables<T> t, const int *__restrict__ rb,
                                                     const int *__restrict__ rl, double *__restrict__ buf,
                                                     int nrecv) {
  const int gid = blockIdx.x * TPB + threadIdx.x;
  if (gid >= NV * nrecv) return;
  const int v = gid / nrecv, r = gid - v * nrecv;
  const int b = rb[r], p = rl[r];
  if (p < 0 || p >= t.nn[b]) {
    buf[gid] = 0.0;
    return;
  }
  T *lam = t.lam[b] + (size_t)v * t.nn[b] + p;
  buf[gid] = (double)*lam;
  *lam = (T)0;
}

// Overset transpose, part 2: each receiver's adjoint goes to its donors
// with the normalised weights (the forward's donors outside the block were
// skipped, and are here).  Receivers sharing a donor add atomically.
template <typename T>
__global__ void __launch_bounds__(TPB) k_adj_ov_give(AdjOvTables<T> t, const int *__restrict__ db,
                                                     const int *__restrict__ dl, const int *__restrict__ w,
                                                     const double *__restrict__ coef, int stride,
                                                     const double *__restrict__ buf, int nrecv) {
  const int gid = blockIdx.x * TPB + threadIdx.x;
  if (gid >= NV * nrecv) return;
  const int v = gid / nrecv, r = gid - v * nrecv;
  const double lr = buf[gid];
  if (lr == 0.0) return;
  const int b = db[r];
  const int ni = t.ni[b], n = t.nn[b];
  const int w1 = w[3 * b], w2 = w[3 * b + 1];
  const double *c = coef + (size_t)r * stride;
  T *lam = t.lam[b] + (size_t)v * n;
  int q = 0;
  for (int b2 = 0; b2 < w2; b2++)
    for (int b1 = 0; b1 < w1; b1++, q++) {
      const int l1 = dl[r] + b1 + b2 * ni;
      if (l1 >= 0 && l1 < n) atomicAdd(&lam[l1], (T)(c[q] * lr));
    }
}

struct AdjBcArgs {
  int type, dir, end;
  int i0, i1, j0, j1;   // face box (collapsed along dir)
  int ni, nj;
  double neu[9];        // closure row 0 of the FD operator
  int sb;
};

// Transpose of one BC face (bc::apply_BC, boundary_condition.cpp:46-209):
// at a face point that is not a hole, the variables a Neumann closure
// (bc:348-444) sets hand their adjoint to the points along the normal
// (weights -neu[s] / neu[0]); wall-slip velocities (bc:448-494) take the
// transposed 2 x 2 map J Minv Z M; dirichlet_allzero and a centerline's
// normal velocity (bc:498-545) pass nothing on.  Points of one face reach
// distinct points along their normals, so no two threads touch one value.
template <typename T>
__global__ void __launch_bounds__(TPB) k_adj_bc_face(T *__restrict__ lam, const double *__restrict__ met,
                                                     const double *__restrict__ jac,
                                                     const int8_t *__restrict__ iblank, AdjBcArgs a) {
  const int fw = a.i1 - a.i0 + 1, fh = a.j1 - a.j0 + 1;
  const int q = blockIdx.x * TPB + threadIdx.x;
  if (q >= fw * fh) return;
  const int i = a.i0 + q % fw, j = a.j0 + q / fw;
  const int n = a.ni * a.nj;
  const int p = j * a.ni + i;
  if (iblank && iblank[p] == BLANKED) return;
  const int mf = a.dir == 0 ? 1 : a.ni;
  const int c = a.dir == 0 ? i : j, len = a.dir == 0 ? a.ni : a.nj;
  const int sg = a.end == 0 ? 1 : -1;
  double lv[NV];
  for (int v = 0; v < NV; v++) lv[v] = (double)lam[(size_t)v * n + p];
  bool neu[NV] = {false, false, false, false, false};
  double keep[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};  // adjoint left at the face point
  switch (a.type) {
    case LEE_BC_NEUMANN:
      for (int v = 0; v < NV; v++) neu[v] = true;
      break;
    case LEE_BC_WALL_SLIP_KINEMATIC: {
      neu[IS] = neu[IP] = true;
      const double m00 = met[p], m01 = met[(size_t)n + p], m10 = met[(size_t)2 * n + p], m11 = met[(size_t)3 * n + p];
      const double J = jac[p];
      // forward: k = M u, k[dir] = 0, u' = J Minv k with Minv = [[m11, -m01], [-m10, m00]]
      double g0 = J * (m11 * lv[IUX] - m10 * lv[IUY]);
      double g1 = J * (-m01 * lv[IUX] + m00 * lv[IUY]);
      if (a.dir == 0) g0 = 0.0; else g1 = 0.0;
      keep[IUX] = m00 * g0 + m10 * g1;
      keep[IUY] = m01 * g0 + m11 * g1;
      keep[IUZ] = lv[IUZ];
      break;
    }
    case LEE_BC_CENTERLINE_CART_NORM2X:
    case LEE_BC_CENTERLINE_CART_NORM2Y:
    case LEE_BC_CENTERLINE_CART_NORM2Z:
      for (int v = 0; v < NV; v++) neu[v] = true;
      neu[IUX + (a.type - LEE_BC_CENTERLINE_CART_NORM2X)] = false;
      break;
    case LEE_BC_CENTERLINE_AXISYM:  // u_r = u_y' set to 0 on the axis (bc:498-545)
      for (int v = 0; v < NV; v++) neu[v] = true;
      neu[IUY] = false;
      break;
    default:  // dirichlet_allzero, and dirichlet_harmonicwave (its data do not depend on the state)
      break;
  }
  for (int v = 0; v < NV; v++) lam[(size_t)v * n + p] = (T)keep[v];
  for (int v = 0; v < NV; v++) {
    if (!neu[v] || lv[v] == 0.0) continue;
    for (int s = 1; s < a.sb && s <= 8; s++) {
      const int cs = c + sg * s;
      if (cs < 0 || cs >= len) continue;
      T *t = &lam[(size_t)v * n + p + sg * s * mf];
      *t = (T)((double)*t - a.neu[s] / a.neu[0] * lv[v]);
    }
  }
}

}  // namespace

// ------------------------------------------------------------- launchers

template <typename T>
static hipError_t adj_stage_t(hipStream_t s, lee_handle *h, Block &b, const AdjStage &st) {
  AdjArgs<T> a{};
  const int R = b.adj_r == 1 ? 1 : 2;  // 8 flux fields of (4R + 4) x 68 points stay within 64 KB of LDS
  a.ni = b.ni;
  a.nj = b.nj;
  a.ntx = (b.ni + ST_TX - 1) / ST_TX;
  a.nty = (b.nj + 4 * R - 1) / (4 * R);
  a.ntiles = a.ntx * a.nty;
  a.chunk = (a.ntiles + 7) / 8;
  T *const *L = sizeof(T) == 8 ? (T *const *)b.A64 : (T *const *)b.A32;
  auto buf = [&](int k) -> T * { return k < 0 ? nullptr : L[k]; };
  a.lacc = buf(st.lacc);
  a.lin = buf(st.lin);
  a.lyin = buf(st.lyin);
  a.lout = buf(st.lout);
  a.lyout = buf(st.lyout);
  a.mu = sizeof(T) == 8 ? (T *)b.amu64 : (T *)b.amu32;
  a.grad = st.grad ? b.agrad : nullptr;
  a.met = b.met;
  a.jac = b.jac;
  a.mean = b.mean;
  a.mgrad = b.mgrad;
  a.rhob = b.rhobar;
  a.u0 = b.umean[IUX];
  a.v0 = b.umean[IUY];
  a.gamma = h->cfg.gamma;
  a.gi = 1.0 / h->cfg.gamma;
  a.gp0 = h->cfg.gamma * b.umean[IP];
  a.rhob0 = std::pow(h->cfg.gamma * b.umean[IP], 1.0 / h->cfg.gamma) * std::exp(-b.umean[IS]);
  a.nz = 0;
  for (const auto &z : b.zones) {
    if (a.nz >= MAX_ZONES) break;
    ZoneT<T> &q = a.z[a.nz++];
    q.i0 = z.is[0];
    q.i1 = z.ie[0];
    q.j0 = z.is[1];
    q.j1 = z.ie[1];
    q.nzi = z.nz[0];
    q.sig = sizeof(T) == 8 ? (const T *)z.sigma64 : (const T *)z.sigma32;
  }
  a.hb = st.hb;
  a.ha = st.ha;
  a.src = st.src;
  a.iblank = b.iblank;
  a.code = b.code_fd;
  a.clean = b.clean_fd;
  if (h->cfg.axisym) {  // an axisymmetric block always has its coefficient fields
    const T *mc = sizeof(T) == 8 ? (const T *)b.mc64 : (const T *)b.mc32;
    a.rad = mc + (size_t)MC_R * b.n;
  }
  const dim3 grid(a.chunk * 8), block(64, 4);
  const bool mf = !b.uniform_mean, ho = b.has_holes;
#define LEE_ADJ(MF, K, HO)                                                                     \
  do {                                                                                         \
    if (R == 1) hipLaunchKernelGGL((k_adj_stage<T, MF, K, 1, HO>), grid, block, 0, s, a);      \
    else hipLaunchKernelGGL((k_adj_stage<T, MF, K, 2, HO>), grid, block, 0, s, a);             \
  } while (0)
#define LEE_ADJK(MF, HO)                                          \
  do {                                                            \
    if (st.kind == STAGE_FIRST) LEE_ADJ(MF, STAGE_FIRST, HO);     \
    else if (st.kind == STAGE_MID) LEE_ADJ(MF, STAGE_MID, HO);    \
    else LEE_ADJ(MF, STAGE_LAST, HO);                             \
  } while (0)
  if (mf && ho) LEE_ADJK(true, true);
  else if (mf) LEE_ADJK(true, false);
  else if (ho) LEE_ADJK(false, true);
  else LEE_ADJK(false, false);
#undef LEE_ADJK
#undef LEE_ADJ
  return hipGetLastError();
}

hipError_t adjoint_stage(hipStream_t s, lee_handle *h, Block &b, const AdjStage &st) {
  if (h->cfg.precision == LEE_FP32) return adj_stage_t<float>(s, h, b, st);
  return adj_stage_t<double>(s, h, b, st);
}

template <typename T>
static hipError_t adj_filter_t(hipStream_t s, lee_handle *h, Block &b, int src, int dst, bool with_mu) {
  AdjFilterArgs<T> a{};
  const int TY = sizeof(T) == 8 ? 4 : 8;
  a.ni = b.ni;
  a.nj = b.nj;
  a.ntx = (b.ni + FL_TX - 1) / FL_TX;
  a.nty = (b.nj + TY - 1) / TY;
  a.ntiles = a.ntx * a.nty;
  a.chunk = (a.ntiles + 7) / 8;
  T *const *L = sizeof(T) == 8 ? (T *const *)b.A64 : (T *const *)b.A32;
  a.src = L[src];
  a.dst = L[dst];
  a.mu = with_mu ? (sizeof(T) == 8 ? (const T *)b.amu64 : (const T *)b.amu32) : nullptr;
  a.mean = b.mean;
  a.rhob = b.rhobar;
  a.gi = 1.0 / h->cfg.gamma;
  // without a filter the step's output is the accumulator itself
  const double blend = h->cfg.filter_order > 0 ? h->cfg.filter_blend : 0.0;
  a.blend = (T)blend;
  a.rem = (T)(1.0 - blend);
  a.iblank = b.iblank;
  a.code = b.code_flt;
  a.clean = b.clean_flt;
  const dim3 grid(a.chunk * 8), block(64, 4);
  // without a filter there are no filter codes: the hole reset alone
  // (blend 0 takes the identity part only)
  const bool ho = b.has_holes;
  if (h->cfg.filter_order == 8) {
    if (ho) hipLaunchKernelGGL((k_adj_filter<T, 8, TY, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_adj_filter<T, 8, TY, false>), grid, block, 0, s, a);
  } else {
    if (ho) hipLaunchKernelGGL((k_adj_filter<T, 4, TY, true>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_adj_filter<T, 4, TY, false>), grid, block, 0, s, a);
  }
  return hipGetLastError();
}

hipError_t adjoint_filter(hipStream_t s, lee_handle *h, Block &b, int src, int dst, bool with_mu) {
  if (h->cfg.precision == LEE_FP32) return adj_filter_t<float>(s, h, b, src, dst, with_mu);
  return adj_filter_t<double>(s, h, b, src, dst, with_mu);
}

hipError_t adjoint_zero(hipStream_t s, lee_handle *h, Block &b, int buf) {
  const size_t cnt = (size_t)NV * b.n;
  if (h->cfg.precision == LEE_FP32) {
    float *p = buf < 0 ? b.amu32 : b.A32[buf];
    hipLaunchKernelGGL(k_fill_t<float>, dim3(nblk(buf < 0 ? b.n : cnt)), dim3(TPB), 0, s, p, buf < 0 ? (size_t)b.n : cnt);
  } else {
    double *p = buf < 0 ? b.amu64 : b.A64[buf];
    hipLaunchKernelGGL(k_fill_t<double>, dim3(nblk(buf < 0 ? b.n : cnt)), dim3(TPB), 0, s, p, buf < 0 ? (size_t)b.n : cnt);
  }
  return hipGetLastError();
}

hipError_t adjoint_add_rt_mu(hipStream_t s, lee_handle *h, Block &b, int buf) {
  if (h->cfg.precision == LEE_FP32)
    hipLaunchKernelGGL(k_add_rt_mu<float>, dim3(nblk(b.n)), dim3(TPB), 0, s, b.A32[buf], b.amu32, b.mean, b.rhobar,
                       1.0 / h->cfg.gamma, b.n);
  else
    hipLaunchKernelGGL(k_add_rt_mu<double>, dim3(nblk(b.n)), dim3(TPB), 0, s, b.A64[buf], b.amu64, b.mean, b.rhobar,
                       1.0 / h->cfg.gamma, b.n);
  return hipGetLastError();
}

hipError_t adjoint_zero_grad(hipStream_t s, Block &b) {
  hipLaunchKernelGGL(k_fill_t<double>, dim3(nblk(b.n)), dim3(TPB), 0, s, b.agrad, (size_t)b.n);
  return hipGetLastError();
}

template <typename T>
static hipError_t adj_boundary_t(hipStream_t s, lee_handle *h, int buf) {
  const Overset &o = h->ov;
  if (o.nrecv > 0) {
    AdjOvTables<T> t{};
    for (int ib = 0; ib < h->nblk; ib++) {
      t.lam[ib] = sizeof(T) == 8 ? (T *)h->b[ib].A64[buf] : (T *)h->b[ib].A32[buf];
      t.ni[ib] = h->b[ib].ni;
      t.nn[ib] = h->b[ib].n;
    }
    const unsigned g = nblk((size_t)NV * o.nrecv);
    hipLaunchKernelGGL(k_adj_ov_take<T>, dim3(g), dim3(TPB), 0, s, t, o.rb, o.rl, o.buf, o.nrecv);
    hipLaunchKernelGGL(k_adj_ov_give<T>, dim3(g), dim3(TPB), 0, s, t, o.db, o.dl, o.w, o.coef, o.stride, o.buf,
                       o.nrecv);
  }
  for (int ib = 0; ib < h->nblk; ib++) {
    Block &b = h->b[ib];
    T *lam = sizeof(T) == 8 ? (T *)b.A64[buf] : (T *)b.A32[buf];
    for (int k = (int)b.faces.size() - 1; k >= 0; k--) {
      const BcFace &f = b.faces[k];
      AdjBcArgs a{};
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
      for (int q = 0; q < 9; q++) a.neu[q] = h->fd.bnd[0][q];
      const size_t cnt = (size_t)(a.i1 - a.i0 + 1) * (a.j1 - a.j0 + 1);
      hipLaunchKernelGGL(k_adj_bc_face<T>, dim3(nblk(cnt)), dim3(TPB), 0, s, lam, b.met, b.jac,
                         b.has_holes ? b.iblank : nullptr, a);
    }
  }
  return hipGetLastError();
}

hipError_t adjoint_boundary(hipStream_t s, lee_handle *h, int buf) {
  if (h->cfg.precision == LEE_FP32) return adj_boundary_t<float>(s, h, buf);
  return adj_boundary_t<double>(s, h, buf);
}

hipError_t observer_energy(hipStream_t s, lee_handle *h, Block &b, double *out) {
  if (b.nobs == 0) return hipSuccess;
  if (h->cfg.precision == LEE_FP32)
    hipLaunchKernelGGL(k_obs_energy<float>, dim3(1), dim3(1024), 0, s, b.Y32[0], b.obs_l, b.obs_w, b.nobs, b.n, out);
  else
    hipLaunchKernelGGL(k_obs_energy<double>, dim3(1), dim3(1024), 0, s, b.Y64[0], b.obs_l, b.obs_w, b.nobs, b.n, out);
  return hipGetLastError();
}

hipError_t observer_gather(hipStream_t s, lee_handle *h, Block &b, double *out) {
  if (b.nobs == 0) return hipSuccess;
  if (h->cfg.precision == LEE_FP32)
    hipLaunchKernelGGL(k_obs_gather<float>, dim3(nblk(b.nobs)), dim3(TPB), 0, s, b.Y32[0], b.obs_l, b.nobs, b.n, out);
  else
    hipLaunchKernelGGL(k_obs_gather<double>, dim3(nblk(b.nobs)), dim3(TPB), 0, s, b.Y64[0], b.obs_l, b.nobs, b.n, out);
  return hipGetLastError();
}

hipError_t observer_seed(hipStream_t s, lee_handle *h, Block &b, int buf, const double *samp, double scale) {
  if (b.nobs == 0) return hipSuccess;
  if (h->cfg.precision == LEE_FP32)
    hipLaunchKernelGGL(k_obs_seed<float>, dim3(nblk(b.nobs)), dim3(TPB), 0, s, b.A32[buf], b.obs_l, b.obs_w, samp,
                       b.nobs, b.n, scale);
  else
    hipLaunchKernelGGL(k_obs_seed<double>, dim3(nblk(b.nobs)), dim3(TPB), 0, s, b.A64[buf], b.obs_l, b.obs_w, samp,
                       b.nobs, b.n, scale);
  return hipGetLastError();
}

}  // namespace lee
