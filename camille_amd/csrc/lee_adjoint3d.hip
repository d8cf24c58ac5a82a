// Discrete adjoint of the 3-D fused time advance (no reference counterpart).
//
// Same transposition as the 2-D adjoint (lee_adjoint.hip): per reversed RK
// stage, kappa = b_i h lambda_acc + a_i h lambda_in, then
// lambda_u = sum_l D_l^T F_l + point-local terms, with the flux fields
// F_l,v = J sum_d m[l][d] A_d,v built from the coefficients A_d,v of
// d/dx_d in each right-hand-side term:
//   s:    A_d = -u_d kappa_s
//   u_i:  A_d = -u_d kappa_ui - delta_id gamma p kappa_p
//   p:    A_d = -u_d kappa_p - kappa_ud / rho
// The stage kernel marches a 64 x 4 column tile through z like the forward
// one: the xi and eta fluxes of each plane (tile + 2-point halo) are staged
// in LDS, the zeta fluxes of the thread's own column sit in a 5-plane
// register window.  The transposed filter is rem I + blend F_xi^T F_eta^T
// F_zeta^T: a zeta^T column pass into scratch, then the eta^T, xi^T plane
// pass with the blend and R^T mu.
#include <hip/hip_runtime.h>

#include <cmath>

#include "lee_device.h"
#include "lee_internal.h"

namespace lee {

namespace {

constexpr int A3_TY = 4;
constexpr int A3_ZC = 32;

inline unsigned nba(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

template <typename T>
__device__ __forceinline__ T fd_wt3(int code, int d) {
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

// (D^T f) at coordinate c of a line of length len: rows of the neighbours.
template <typename T>
__device__ __forceinline__ T fd_t3(const T *f, int st, int c, int len) {
  if (c >= 4 && c <= len - 5) {  // every neighbour row is interior: D^T = -D
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
    acc += fd_wt3<T>(default_code(j, len, 2), -o) * f[o * st];
  }
  return acc;
}

template <typename T, int ORDER>
__device__ __forceinline__ T flt_wt3(int code, int d) {
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
__device__ __forceinline__ T flt_t3(const T *f, int st, int c, int len) {
  constexpr int H = ORDER / 2;
  if (c >= 2 * H && c <= len - 1 - 2 * H) return filter_row<T, ORDER>(f, st, CODE_INTERIOR);
  T acc = 0;
#pragma unroll
  for (int o = -H; o <= H; o++) {
    const int j = c + o;
    if (j < 0 || j >= len) continue;
    acc += flt_wt3<T, ORDER>(default_code(j, len, H), -o) * f[o * st];
  }
  return acc;
}

template <typename T>
struct Zone3A {
  int i0, i1, j0, j1, k0, k1, nzi, nzj;
  const T *sig;
};

template <typename T>
struct Adj3Args {
  int ni, nj, nk, ntx, nty, nzc;
  const T *lacc, *lin, *lyin;
  T *lout, *lyout, *mu;
  double *grad;
  const double *met, *jac;
  const T *mc;               // MC3_COUNT coefficient fields (non-uniform mean)
  double u0, v0, w0, gp0, rhob0, gamma;
  int nz;
  Zone3A<T> z[MAX_ZONES];
  double hb, ha, src;
};

// kappa and the three flux vectors (5 variables) at point p of plane k.
template <typename T, bool MEANF, int KIND>
__device__ __forceinline__ void adj_flux3(const Adj3Args<T> &a, size_t p, size_t n, T kap[5], T F[3][5]) {
  const T hb = (T)a.hb, ha = (T)a.ha;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    T kv = hb * a.lacc[v * n + p];
    if (KIND != STAGE_LAST && a.lin) kv += ha * a.lin[v * n + p];
    kap[v] = kv;
  }
  double ub[3] = {a.u0, a.v0, a.w0}, gp = a.gp0, irb = 1.0 / a.rhob0;
  if (MEANF) {
    ub[0] = a.mc[MC3_U * n + p];
    ub[1] = a.mc[MC3_V * n + p];
    ub[2] = a.mc[MC3_W * n + p];
    gp = a.mc[MC3_GP * n + p];
    irb = a.mc[MC3_IRB * n + p];
  }
  double A[3][5];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    A[d][IS] = -ub[d] * (double)kap[IS];
#pragma unroll
    for (int i = 0; i < 3; i++) A[d][IUX + i] = -ub[d] * (double)kap[IUX + i] - (i == d ? gp * (double)kap[IP] : 0.0);
    A[d][IP] = -ub[d] * (double)kap[IP] - (double)kap[IUX + d] * irb;
  }
  double m[3][3];
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int i = 0; i < 3; i++) m[l][i] = a.met[(size_t)(3 * l + i) * n + p];
  const double J = a.jac[p];
#pragma unroll
  for (int l = 0; l < 3; l++)
#pragma unroll
    for (int v = 0; v < 5; v++) {
      double f = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) f += m[l][d] * A[d][v];
      F[l][v] = (T)(f * J);
    }
}

template <typename T, bool MEANF, int KIND>
__global__ void __launch_bounds__(TPB) k_adj_stage3d(Adj3Args<T> a) {
  constexpr int RS = ST_TX + 2 * ST_H, NR = A3_TY + 2 * ST_H;
  __shared__ T S[10][NR][RS];  // xi fluxes then eta fluxes of the 5 variables
  // XCD-banded tiles within each z chunk, as the forward stage
  const int ntl = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, (ntl + 7) / 8);
  if (tile >= ntl) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * ST_TX, j0 = (tile / a.ntx) * A3_TY;
  const int k0 = blockIdx.y * A3_ZC, k1 = min(k0 + A3_ZC, a.nk);
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const size_t plane = (size_t)ni * nj, n = plane * nk;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool own = gi < ni && gj < nj;
  const size_t col = own ? (size_t)gj * ni + gi : 0;
  // zeta fluxes of the own column, planes k - 2 .. k + 2
  T wz[5][5];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int kk = k0 - 2 + q;
    T kap[5], F[3][5];
    if (own && kk >= 0 && kk < nk) adj_flux3<T, MEANF, KIND>(a, (size_t)kk * plane + col, n, kap, F);
    else
#pragma unroll
      for (int v = 0; v < 5; v++) F[2][v] = (T)0;
#pragma unroll
    for (int v = 0; v < 5; v++) wz[v][q] = F[2][v];
  }
  for (int k = k0; k < k1; k++) {
    {
      const int kk = k + 2;
      T kap[5], F[3][5];
      if (own && kk < nk) adj_flux3<T, MEANF, KIND>(a, (size_t)kk * plane + col, n, kap, F);
      else
#pragma unroll
        for (int v = 0; v < 5; v++) F[2][v] = (T)0;
#pragma unroll
      for (int v = 0; v < 5; v++) wz[v][4] = F[2][v];
    }
    // xi and eta fluxes of plane k over the tile and its halo
#pragma unroll
    for (int mm = 0; mm < 2; mm++) {
      const int r = ty + 4 * mm;
      const int rj = j0 - ST_H + r;
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const int c = half == 0 ? tx : ST_TX + tx;
        if (half == 1 && tx >= 2 * ST_H) continue;
        const int ri = i0 - ST_H + c;
        const bool ok = rj >= 0 && rj < nj && ri >= 0 && ri < ni;
        T kap[5], F[3][5];
        if (ok) adj_flux3<T, MEANF, KIND>(a, (size_t)k * plane + (size_t)rj * ni + ri, n, kap, F);
#pragma unroll
        for (int v = 0; v < 5; v++) {
          S[v][r][c] = ok ? F[0][v] : (T)0;
          S[5 + v][r][c] = ok ? F[1][v] : (T)0;
        }
      }
    }
    __syncthreads();
    if (own) {
      const size_t p = (size_t)k * plane + col;
      T lam[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        lam[v] = fd_t3<T>(&S[v][ty + ST_H][tx + ST_H], 1, gi, ni);
        lam[v] += fd_t3<T>(&S[5 + v][ty + ST_H][tx + ST_H], RS, gj, nj);
        lam[v] += fd_t3<T>(&wz[v][2], 1, k, nk);
      }
      T kap[5], lacc[5], lin[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        lacc[v] = a.lacc[v * n + p];
        lin[v] = (KIND != STAGE_LAST && a.lin) ? a.lin[v * n + p] : (T)0;
        T kv = (T)a.hb * lacc[v];
        if (KIND != STAGE_LAST && a.lin) kv += (T)a.ha * lin[v];
        kap[v] = kv;
      }
#pragma unroll 1
      for (int z = 0; z < a.nz; z++) {
        const Zone3A<T> &zz = a.z[z];
        if (gi >= zz.i0 && gi <= zz.i1 && gj >= zz.j0 && gj <= zz.j1 && k >= zz.k0 && k <= zz.k1) {
          const T sg = zz.sig[((size_t)(k - zz.k0) * zz.nzj + (gj - zz.j0)) * zz.nzi + (gi - zz.i0)];
#pragma unroll
          for (int v = 0; v < 5; v++) lam[v] += -sg * kap[v];
        }
      }
      T mur = 0;
      if (MEANF) {
        const T *mc = a.mc;
        auto G = [&](int v, int d) { return mc[(MC3_G + 3 * v + d) * n + p]; };
        const T ub[3] = {mc[MC3_U * n + p], mc[MC3_V * n + p], mc[MC3_W * n + p]};
        const T g = (T)a.gamma, irb = mc[MC3_IRB * n + p];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          T acc = G(IS, d) * kap[IS];
#pragma unroll
          for (int i = 0; i < 3; i++) acc += G(IUX + i, d) * kap[IUX + i];
          acc += G(IP, d) * kap[IP];
          lam[IUX + d] -= acc;
        }
        T div = 0;
#pragma unroll
        for (int d = 0; d < 3; d++) div += G(IUX + d, d);
        lam[IP] -= g * div * kap[IP];
        T mu = 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
          T s = 0;
#pragma unroll
          for (int d = 0; d < 3; d++) s += ub[d] * G(IUX + i, d);
          mu += s * kap[IUX + i];
        }
        mur = -(mu * irb);
        if (KIND != STAGE_FIRST) {
          lam[IS] += -(mc[MC3_RB * n + p] * mur);
          lam[IP] += mc[MC3_RP * n + p] * mur;
        }
      }
      if (a.grad) a.grad[p] += a.src * (double)kap[IP];
      if (KIND == STAGE_FIRST) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
          T out = lam[v] + lacc[v] + lin[v];
          if (a.lyin) out += a.lyin[v * n + p];
          a.lout[v * n + p] = out;
        }
        if (MEANF) a.mu[p] = mur;
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) a.lout[v * n + p] = lam[v];
        if (a.lyout) {
#pragma unroll
          for (int v = 0; v < 5; v++) {
            T y = lin[v];
            if (a.lyin) y += a.lyin[v * n + p];
            a.lyout[v * n + p] = y;
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
      for (int q = 0; q < 4; q++) wz[v][q] = wz[v][q + 1];
  }
}

// zeta^T column pass: tmp = F_zeta^T lam.
template <typename T, int ORDER>
__global__ void __launch_bounds__(TPB) k_adj_filter3d_z(const T *__restrict__ lam, T *__restrict__ tmp, int ni,
                                                        int nj, int nk) {
  constexpr int H = ORDER / 2;
  const size_t plane = (size_t)ni * nj, n = plane * nk;
  const size_t col = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (col >= plane) return;
  const int k0 = blockIdx.y * A3_ZC, k1 = min(k0 + A3_ZC, nk);
  T w[5][2 * H + 1];
#pragma unroll
  for (int q = 0; q < 2 * H; q++) {
    const int kk = k0 - H + q;
    const bool ok = kk >= 0 && kk < nk;
#pragma unroll
    for (int v = 0; v < 5; v++) w[v][q] = ok ? lam[v * n + (size_t)kk * plane + col] : (T)0;
  }
  // four planes per pass, their window planes requested together (clamped
  // plane index, zeros selected afterwards)
  auto plane_out = [&](int k, const T (&wn)[5]) {
#pragma unroll
    for (int v = 0; v < 5; v++) w[v][2 * H] = wn[v];
#pragma unroll
    for (int v = 0; v < 5; v++) tmp[v * n + (size_t)k * plane + col] = flt_t3<T, ORDER>(&w[v][H], 1, k, nk);
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
      for (int q = 0; q < 2 * H; q++) w[v][q] = w[v][q + 1];
  };
  int k = k0;
  for (; k + 4 <= k1; k += 4) {
    T wn[4][5];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = k + u + H;
      const size_t pc = (size_t)min(kk, nk - 1) * plane + col;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T t = lam[v * n + pc];
        wn[u][v] = kk < nk ? t : (T)0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) plane_out(k + u, wn[u]);
  }
  for (; k < k1; k++) {
    const int kk = k + H;
    T wn[5];
#pragma unroll
    for (int v = 0; v < 5; v++) wn[v] = kk < nk ? lam[v * n + (size_t)kk * plane + col] : (T)0;
    plane_out(k, wn);
  }
}

template <typename T>
struct AdjF3Args {
  int ni, nj, nk, ntx, nty, ntiles;
  const T *tmp;    // zeta^T-filtered adjoint
  const T *orig;   // adjoint of the filtered state
  T *dst;
  const T *mu;     // null: none
  const T *mc;
  T rem, blend;
};

// eta^T then xi^T sweeps of plane blockIdx.y, blend and R^T mu.
template <typename T, int ORDER, int TY>
__global__ void __launch_bounds__(TPB) k_adj_filter3d_plane(AdjF3Args<T> a) {
  constexpr int NR = TY + 2 * FL_H, NC = FL_TX + 2 * FL_H, KR = NR / 4;
  __shared__ T in[5][NR][NC];
  __shared__ T mid[5][TY][NC];
  const int tile = xcd_tile(blockIdx.x, (a.ntiles + 7) / 8);  // XCD-banded tiles of each plane
  if (tile >= a.ntiles) return;
  const int k = blockIdx.y;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj;
  const size_t plane = (size_t)ni * nj, n = plane * a.nk, off = (size_t)k * plane;
#pragma unroll
  for (int kr = 0; kr < KR; kr++) {
    const int r = ty + 4 * kr;
    const int gj = j0 - FL_H + r;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int c = half == 0 ? tx : FL_TX + tx;
      if (half == 1 && tx >= 2 * FL_H) continue;
      const int gi = i0 - FL_H + c;
      const bool ok = gj >= 0 && gj < nj && gi >= 0 && gi < ni;
#pragma unroll
      for (int v = 0; v < 5; v++) in[v][r][c] = ok ? a.tmp[v * n + off + (size_t)gj * ni + gi] : (T)0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int kr = 0; kr < TY / 4; kr++) {
    const int r = ty + 4 * kr;
    const int gj = j0 + r;
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int c = half == 0 ? tx : FL_TX + tx;
      if (half == 1 && tx >= 2 * FL_H) continue;
      const int gi = i0 - FL_H + c;
      const bool ok = gj < nj && gi >= 0 && gi < ni;
#pragma unroll
      for (int v = 0; v < 5; v++) mid[v][r][c] = ok ? flt_t3<T, ORDER>(&in[v][r + FL_H][c], NC, gj, nj) : (T)0;
    }
  }
  __syncthreads();
  const int gi = i0 + tx;
#pragma unroll
  for (int kr = 0; kr < TY / 4; kr++) {
    const int r = ty + 4 * kr;
    const int gj = j0 + r;
    if (gi >= ni || gj >= nj) continue;
    const size_t p = off + (size_t)gj * ni + gi;
    T out[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const T f = flt_t3<T, ORDER>(&mid[v][r][tx + FL_H], 1, gi, ni);
      out[v] = a.rem * a.orig[v * n + p] + a.blend * f;
    }
    if (a.mu) {
      const T m = a.mu[p];
      out[IS] += -(a.mc[MC3_RB * n + p] * m);
      out[IP] += a.mc[MC3_RP * n + p] * m;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) a.dst[v * n + p] = out[v];
  }
}

template <typename T>
__global__ void __launch_bounds__(TPB) k_add_rt_mu3(T *__restrict__ lam, const T *__restrict__ mu,
                                                    const T *__restrict__ mc, size_t n) {
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  const T m = mu[p];
  lam[IS * n + p] += -(mc[MC3_RB * n + p] * m);
  lam[IP * n + p] += mc[MC3_RP * n + p] * m;
}

}  // namespace

// ------------------------------------------------------------- launchers

template <typename T>
static hipError_t adj_stage3d_t(hipStream_t s, lee_handle *h, Block &b, const AdjStage &st) {
  Adj3Args<T> a{};
  a.ni = b.ni;
  a.nj = b.nj;
  a.nk = b.nk;
  a.ntx = (b.ni + ST_TX - 1) / ST_TX;
  a.nty = (b.nj + A3_TY - 1) / A3_TY;
  a.nzc = (b.nk + A3_ZC - 1) / A3_ZC;
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
  a.mc = sizeof(T) == 8 ? (const T *)b.mc64 : (const T *)b.mc32;
  a.u0 = b.umean[IUX];
  a.v0 = b.umean[IUY];
  a.w0 = b.umean[IUZ];
  a.gamma = h->cfg.gamma;
  a.gp0 = h->cfg.gamma * b.umean[IP];
  a.rhob0 = std::pow(h->cfg.gamma * b.umean[IP], 1.0 / h->cfg.gamma) * std::exp(-b.umean[IS]);
  a.nz = 0;
  for (const auto &z : b.zones) {
    if (a.nz >= MAX_ZONES) break;
    Zone3A<T> &q = a.z[a.nz++];
    q.i0 = z.is[0];
    q.i1 = z.ie[0];
    q.j0 = z.is[1];
    q.j1 = z.ie[1];
    q.k0 = z.is[2];
    q.k1 = z.ie[2];
    q.nzi = z.nz[0];
    q.nzj = z.nz[1];
    q.sig = sizeof(T) == 8 ? (const T *)z.sigma64 : (const T *)z.sigma32;
  }
  a.hb = st.hb;
  a.ha = st.ha;
  a.src = st.src;
  const dim3 grid((a.ntx * a.nty + 7) / 8 * 8, a.nzc), block(ST_TX, A3_TY);
  const bool mf = !b.uniform_mean;
#define LEE_A3(MF)                                                                                            \
  do {                                                                                                        \
    if (st.kind == STAGE_FIRST) hipLaunchKernelGGL((k_adj_stage3d<T, MF, STAGE_FIRST>), grid, block, 0, s, a); \
    else if (st.kind == STAGE_MID) hipLaunchKernelGGL((k_adj_stage3d<T, MF, STAGE_MID>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_adj_stage3d<T, MF, STAGE_LAST>), grid, block, 0, s, a);                         \
  } while (0)
  if (mf) LEE_A3(true);
  else LEE_A3(false);
#undef LEE_A3
  return hipGetLastError();
}

hipError_t adjoint_stage3d(hipStream_t s, lee_handle *h, Block &b, const AdjStage &st) {
  if (h->cfg.precision == LEE_FP32) return adj_stage3d_t<float>(s, h, b, st);
  return adj_stage3d_t<double>(s, h, b, st);
}

// dst = rem src + blend F_xi^T F_eta^T F_zeta^T src (+ R^T mu); scratch
// holds the zeta^T pass (all buffers index Block::A).
template <typename T>
static hipError_t adj_filter3d_t(hipStream_t s, lee_handle *h, Block &b, int src, int dst, int scratch,
                                 bool with_mu) {
  T *const *L = sizeof(T) == 8 ? (T *const *)b.A64 : (T *const *)b.A32;
  const size_t plane = (size_t)b.ni * b.nj;
  const bool o8 = h->cfg.filter_order == 8;
  const dim3 gz(nba(plane), (b.nk + A3_ZC - 1) / A3_ZC);
  if (o8) hipLaunchKernelGGL((k_adj_filter3d_z<T, 8>), gz, dim3(TPB), 0, s, L[src], L[scratch], b.ni, b.nj, b.nk);
  else hipLaunchKernelGGL((k_adj_filter3d_z<T, 4>), gz, dim3(TPB), 0, s, L[src], L[scratch], b.ni, b.nj, b.nk);
  constexpr int TY = 4;
  AdjF3Args<T> a{};
  a.ni = b.ni;
  a.nj = b.nj;
  a.nk = b.nk;
  a.ntx = (b.ni + FL_TX - 1) / FL_TX;
  a.nty = (b.nj + TY - 1) / TY;
  a.ntiles = a.ntx * a.nty;
  a.tmp = L[scratch];
  a.orig = L[src];
  a.dst = L[dst];
  a.mu = with_mu ? (sizeof(T) == 8 ? (const T *)b.amu64 : (const T *)b.amu32) : nullptr;
  a.mc = sizeof(T) == 8 ? (const T *)b.mc64 : (const T *)b.mc32;
  const double blend = h->cfg.filter_order > 0 ? h->cfg.filter_blend : 0.0;
  a.blend = (T)blend;
  a.rem = (T)(1.0 - blend);
  const dim3 grid((a.ntiles + 7) / 8 * 8, b.nk), block(64, 4);
  if (o8) hipLaunchKernelGGL((k_adj_filter3d_plane<T, 8, TY>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_adj_filter3d_plane<T, 4, TY>), grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t adjoint_filter3d(hipStream_t s, lee_handle *h, Block &b, int src, int dst, int scratch, bool with_mu) {
  if (h->cfg.precision == LEE_FP32) return adj_filter3d_t<float>(s, h, b, src, dst, scratch, with_mu);
  return adj_filter3d_t<double>(s, h, b, src, dst, scratch, with_mu);
}

hipError_t adjoint_add_rt_mu3(hipStream_t s, lee_handle *h, Block &b, int buf) {
  if (h->cfg.precision == LEE_FP32)
    hipLaunchKernelGGL(k_add_rt_mu3<float>, dim3(nba(b.n)), dim3(TPB), 0, s, b.A32[buf], b.amu32, b.mc32, (size_t)b.n);
  else
    hipLaunchKernelGGL(k_add_rt_mu3<double>, dim3(nba(b.n)), dim3(TPB), 0, s, b.A64[buf], b.amu64, b.mc64,
                       (size_t)b.n);
  return hipGetLastError();
}

}  // namespace lee
