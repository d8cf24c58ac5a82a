// 3-D kernels of the fused LEE time advance (gfx950).
//
// The reference computes 3-D metrics but stops at the 3-D Jacobian
// (metrics.cpp:292), so it never runs a 3-D case; these kernels follow
// its 3-D operator definitions (take_derivative_xyz over xi, eta, zeta,
// spatial:1802-1826; the LEE right-hand side summed over three directions,
// spatial:2053-2130; filter sweeps xi, eta, zeta in turn, spatial:2191-2246)
// and are checked against the 2-D path on z-invariant problems and by axis
// permutations.
//
// Stage kernel: a 64 x 4 tile of (i, j) columns marches through a chunk of
// z planes.  Each plane's tile plus its 2-point (i, j) halo is staged in
// LDS for the xi and eta derivatives; the zeta derivative comes from a
// 5-plane register window of the thread's own column.  Filter: one pass of
// the 2-D xi-eta tile filter per plane, then a column pass in zeta with
// the blend.
#include <hip/hip_runtime.h>

#include <cmath>

#include "lee_device.h"
#include "lee_internal.h"

namespace lee {

namespace {

constexpr int S3_TY = 4;      // tile rows in j
#ifndef LEE_Z_CHUNK
#define LEE_Z_CHUNK 32
#endif
constexpr int Z_CHUNK = LEE_Z_CHUNK;  // planes per z chunk
constexpr int FUSED_Z_CHUNK = 64;     // planes per chunk of the single-pass filter (4 halo planes a side)

inline unsigned nb3(size_t n) { return (unsigned)((n + TPB - 1) / TPB); }

// ---------------------------------------------------------------- set-up

// FD derivative along index direction dir (0 xi, 1 eta, 2 zeta), rows of
// StandardCentral OA4 with the block-edge closures.
__global__ void __launch_bounds__(TPB) k_op_dir3(const double *__restrict__ f, double *__restrict__ df, int ni,
                                                 int nj, int nk, int dir) {
  const size_t n = (size_t)ni * nj * nk;
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p % ni), j = (int)((p / ni) % nj), k = (int)(p / ((size_t)ni * nj));
  const int c = dir == 0 ? i : (dir == 1 ? j : k), len = dir == 0 ? ni : (dir == 1 ? nj : nk);
  const int st = dir == 0 ? 1 : (dir == 1 ? ni : ni * nj);
  df[p] = fd_row<double>(f + p, st, default_code(c, len, 2));
}

// m[l][i] = (1/J) d xi_l / d x_i as the adjugate of A = d x_i / d xi_l,
// J = 1 / det A.  xx holds A row-major: xx[(3 i + l) n + p].
__global__ void __launch_bounds__(TPB) k_metrics3d(const double *__restrict__ xx, double *__restrict__ met,
                                                   double *__restrict__ jac, size_t n) {
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  double A[3][3];
  for (int i = 0; i < 3; i++)
    for (int l = 0; l < 3; l++) A[i][l] = xx[(size_t)(3 * i + l) * n + p];
  double m[3][3];
  m[0][0] = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  m[0][1] = -(A[0][1] * A[2][2] - A[0][2] * A[2][1]);
  m[0][2] = A[0][1] * A[1][2] - A[0][2] * A[1][1];
  m[1][0] = -(A[1][0] * A[2][2] - A[1][2] * A[2][0]);
  m[1][1] = A[0][0] * A[2][2] - A[0][2] * A[2][0];
  m[1][2] = -(A[0][0] * A[1][2] - A[0][2] * A[1][0]);
  m[2][0] = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  m[2][1] = -(A[0][0] * A[2][1] - A[0][1] * A[2][0]);
  m[2][2] = A[0][0] * A[1][1] - A[0][1] * A[1][0];
  for (int l = 0; l < 3; l++)
    for (int i = 0; i < 3; i++) met[(size_t)(3 * l + i) * n + p] = m[l][i];
  const double det = A[0][0] * m[0][0] + A[0][1] * m[1][0] + A[0][2] * m[2][0];
  jac[p] = 1.0 / det;
}

// take_derivative_xyz in 3-D: out = (sum_l m[l][d] d f / d xi_l) * J.
__global__ void __launch_bounds__(TPB) k_deriv_xyz3(const double *__restrict__ f, double *__restrict__ out,
                                                    const double *__restrict__ met, const double *__restrict__ jac,
                                                    int ni, int nj, int nk, int d) {
  const size_t n = (size_t)ni * nj * nk;
  const size_t p = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p % ni), j = (int)((p / ni) % nj), k = (int)(p / ((size_t)ni * nj));
  const double d0 = fd_row<double>(f + p, 1, default_code(i, ni, 2));
  const double d1 = fd_row<double>(f + p, ni, default_code(j, nj, 2));
  const double d2 = fd_row<double>(f + p, ni * nj, default_code(k, nk, 2));
  double sum = 0.0;
  sum += met[(size_t)(0 + d) * n + p] * d0;
  sum += met[(size_t)(3 + d) * n + p] * d1;
  sum += met[(size_t)(6 + d) * n + p] * d2;
  out[p] = sum * jac[p];
}

// ------------------------------------------------------------ stage (3-D)

template <typename T>
struct Zone3 {
  int i0, i1, j0, j1, k0, k1, nzi, nzj;
  const T *sig;
};

template <typename T>
struct Stage3Args {
  int ni, nj, nk, ntx, nty, nzc;
  const T *yin, *yn;
  T *acc, *yout;
  const double *met, *jac;
  const T *mc;                      // non-uniform mean: MC3_COUNT fields
  const T *rho_in;
  T *rho_out;
  double u0, v0, w0, gp0, rhob0, gamma;
  int nz;
  Zone3<T> z[MAX_ZONES];
  int si0, si1, sj0, sj1, sk0, sk1;
  const double *g;
  double s1, s2;
  T h, b, a;
};

// One RK stage on a 64 x 4 column tile through a chunk of Z_CHUNK planes.
template <typename T, bool MEANF, int KIND>
__global__ void __launch_bounds__(TPB) k_stage3d(Stage3Args<T> a) {
  constexpr int RS = ST_TX + 2 * ST_H, NR = S3_TY + 2 * ST_H;
  __shared__ T s[5][NR][RS];
  // within each z chunk, XCD x (workgroups x mod 8 of a grid row that is a
  // multiple of 8 wide) takes the x-th band of tiles, so that the halo a
  // tile reads of its neighbours is mostly in the same XCD's L2
  const int ntl = a.ntx * a.nty;
  const int tile = xcd_tile(blockIdx.x, (ntl + 7) / 8);
  if (tile >= ntl) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * ST_TX, j0 = (tile / a.ntx) * S3_TY;
  const int k0 = blockIdx.y * Z_CHUNK, k1 = min(k0 + Z_CHUNK, a.nk);
  const int ni = a.ni, nj = a.nj, nk = a.nk;
  const size_t plane = (size_t)ni * nj, n = plane * nk;
  const int gi = i0 + tx, gj = j0 + ty;
  const bool own = gi < ni && gj < nj;
  const size_t col = own ? (size_t)gj * ni + gi : 0;
  // register window of the own column: planes k - 2 .. k + 2
  T w[5][5];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int kk = k0 - 2 + q;
    const bool ok = own && kk >= 0 && kk < nk;
#pragma unroll
    for (int v = 0; v < 5; v++) w[v][q] = ok ? a.yin[v * n + (size_t)kk * plane + col] : (T)0;
  }
  for (int k = k0; k < k1; k++) {
    // every global operand of plane k is requested before the barrier, so
    // one memory latency per plane is exposed instead of three
    {
      const int kk = k + 2;
      const bool ok = own && kk < nk;
#pragma unroll
      for (int v = 0; v < 5; v++) w[v][4] = ok ? a.yin[v * n + (size_t)kk * plane + col] : (T)0;
    }
    // plane k with its (i, j) halo.  The centre (region rows 2..5, columns
    // 2..65) is the window's plane k, already in registers; only the ring
    // is loaded: region row hr (rows 0, 1 for ty 0, 1; rows 6, 7 for ty 2,
    // 3) at columns tx and 64 + tx, and for tx < 4 the side column sc (0,
    // 1, 66, 67) of region row 2 + ty
    const int hr = ty < 2 ? ty : ty + 4, sc = tx < 2 ? tx : ST_TX + tx;
    T th[2][5], ts[5];
    {
      const int rj = j0 - ST_H + hr;
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const int c = half == 0 ? tx : ST_TX + tx;
        const int ri = i0 - ST_H + c;
        const bool ok = (half == 0 || tx < 2 * ST_H) && rj >= 0 && rj < nj && ri >= 0 && ri < ni;
        const size_t q = (size_t)k * plane + (size_t)(ok ? rj : 0) * ni + (ok ? ri : 0);
#pragma unroll
        for (int v = 0; v < 5; v++) th[half][v] = ok ? a.yin[v * n + q] : (T)0;
      }
      const int sj = j0 + ty, si = i0 - ST_H + sc;
      const bool ok = tx < 2 * ST_H && sj < nj && si >= 0 && si < ni;
      const size_t q = (size_t)k * plane + (size_t)(ok ? sj : 0) * ni + (ok ? si : 0);
#pragma unroll
      for (int v = 0; v < 5; v++) ts[v] = ok ? a.yin[v * n + q] : (T)0;
    }
    // per-point operands (a clamped column for threads outside the block)
    const size_t p = (size_t)k * plane + col;
    double m[3][3];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int i = 0; i < 3; i++) m[l][i] = a.met[(size_t)(3 * l + i) * n + p];
    const double J = a.jac[p];
    T mcv[MC3_COUNT];
    if (MEANF) {
#pragma unroll
      for (int f = 0; f < MC3_COUNT; f++)
        mcv[f] = (f < MC3_RP || KIND != STAGE_FIRST) ? a.mc[(size_t)f * n + p] : (T)0;
    }
    const T rho_first = (MEANF && KIND == STAGE_FIRST) ? a.rho_in[p] : (T)0;
    T accp[5], ynp[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      accp[v] = KIND != STAGE_FIRST ? a.acc[v * n + p] : (T)0;
      ynp[v] = KIND == STAGE_MID ? a.yn[v * n + p] : (T)0;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
      s[v][hr][tx] = th[0][v];
      s[v][ty + ST_H][tx + ST_H] = w[v][2];
      if (tx < 2 * ST_H) {
        s[v][hr][ST_TX + tx] = th[1][v];
        s[v][ty + ST_H][sc] = ts[v];
      }
    }
    __syncthreads();
    if (own) {
      const int cx = default_code(gi, ni, 2), cy = default_code(gj, nj, 2), cz = default_code(k, nk, 2);
      T d[5][3], yc[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T *sp = &s[v][ty + ST_H][tx + ST_H];
        yc[v] = w[v][2];
        d[v][0] = fd_row<T>(sp, 1, cx);
        d[v][1] = fd_row<T>(sp, RS, cy);
        d[v][2] = fd_row<T>(&w[v][2], 1, cz);
      }
      T f[5][3];  // f[v][x_i] = d var / d x_i
#pragma unroll
      for (int v = 0; v < 5; v++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          double acc = 0.0;
          acc += m[0][i] * (double)d[v][0];
          acc += m[1][i] * (double)d[v][1];
          acc += m[2][i] * (double)d[v][2];
          f[v][i] = (T)(acc * J);
        }
      T r[5];
      const T up[3] = {yc[IUX], yc[IUY], yc[IUZ]};
      if (!MEANF) {
        const T ub[3] = {(T)a.u0, (T)a.v0, (T)a.w0};
        const T gp = (T)a.gp0, irb = (T)(1.0 / a.rhob0);
        r[IS] = 0;
        for (int dd = 0; dd < 3; dd++) r[IS] -= ub[dd] * f[IS][dd];
#pragma unroll
        for (int vi = 0; vi < 3; vi++) {
          T rv = 0;
#pragma unroll
          for (int dd = 0; dd < 3; dd++) rv -= ub[dd] * f[IUX + vi][dd];
          rv -= f[IP][vi] * irb;
          r[IUX + vi] = rv;
        }
        T rp = 0;
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
          rp -= ub[dd] * f[IP][dd];
          rp -= gp * f[IUX + dd][dd];
        }
        r[IP] = rp;
      } else {
        const T ub[3] = {mcv[MC3_U], mcv[MC3_V], mcv[MC3_W]};
        const T gp = mcv[MC3_GP], irb = mcv[MC3_IRB];
        auto G = [&](int v, int dd) { return mcv[MC3_G + 3 * v + dd]; };
        T rho;
        if (KIND == STAGE_FIRST) rho = rho_first;
        else rho = yc[IP] * mcv[MC3_RP] - yc[IS] * mcv[MC3_RB];
        const T fr = rho * irb;
        const T g = (T)a.gamma;
        T rs = 0;
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
          rs -= ub[dd] * f[IS][dd];
          rs -= up[dd] * G(IS, dd);
        }
        r[IS] = rs;
#pragma unroll
        for (int vi = 0; vi < 3; vi++) {
          T rv = 0;
#pragma unroll
          for (int dd = 0; dd < 3; dd++) {
            rv -= ub[dd] * f[IUX + vi][dd];
            rv -= (up[dd] + fr * ub[dd]) * G(IUX + vi, dd);
          }
          rv -= f[IP][vi] * irb;
          r[IUX + vi] = rv;
        }
        T rp = 0;
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
          rp -= ub[dd] * f[IP][dd];
          rp -= up[dd] * G(IP, dd);
          rp -= gp * f[IUX + dd][dd];
          rp -= g * yc[IP] * G(IUX + dd, dd);
        }
        r[IP] = rp;
      }
      if (gi >= a.si0 && gi <= a.si1 && gj >= a.sj0 && gj <= a.sj1 && k >= a.sk0 && k <= a.sk1) {
        const double gval =
            a.g[((size_t)(k - a.sk0) * (a.sj1 - a.sj0 + 1) + (gj - a.sj0)) * (a.si1 - a.si0 + 1) + (gi - a.si0)];
        r[IP] += (T)(a.s1 * gval * a.s2);
      }
#pragma unroll 1
      for (int z = 0; z < a.nz; z++) {
        const Zone3<T> &zz = a.z[z];
        if (gi >= zz.i0 && gi <= zz.i1 && gj >= zz.j0 && gj <= zz.j1 && k >= zz.k0 && k <= zz.k1) {
          const T sg = zz.sig[((size_t)(k - zz.k0) * zz.nzj + (gj - zz.j0)) * zz.nzi + (gi - zz.i0)];
#pragma unroll
          for (int v = 0; v < 5; v++) r[v] += -sg * yc[v];
        }
      }
      T accv[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const size_t q = v * n + p;
        const T rh = r[v] * a.h;
        const T acc_in = KIND == STAGE_FIRST ? yc[v] : accp[v];
        accv[v] = acc_in + a.b * rh;
        a.acc[q] = accv[v];
        if (KIND != STAGE_LAST) {
          const T yn = KIND == STAGE_FIRST ? yc[v] : ynp[v];
          a.yout[q] = yn + a.a * rh;
        }
      }
      if (MEANF && KIND == STAGE_LAST) a.rho_out[p] = accv[IP] * mcv[MC3_RP] - accv[IS] * mcv[MC3_RB];
    }
    __syncthreads();  // the next plane overwrites the LDS tile
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
      for (int q = 0; q < 4; q++) w[v][q] = w[v][q + 1];
  }
}

// ------------------------------------------------------------ filter (3-D)

// Pass 1: the xi and eta sweeps of plane blockIdx.y (the 2-D tile filter
// with blend 1) into tmp.
template <typename T, int ORDER, int TY>
__global__ void __launch_bounds__(TPB) k_filter3d_plane(FilterArgs<T> a, size_t plane) {
  __shared__ FilterLds<T, 5, TY> L;
  const int tile = xcd_tile(blockIdx.x, a.chunk);
  if (tile >= a.ntiles) return;
  FilterArgs<T> b = a;
  b.src = a.src + (size_t)blockIdx.y * plane;
  b.dst = a.dst + (size_t)blockIdx.y * plane;
  filter_tile<T, false, ORDER, 5, TY>(b, tile, L);
}

// Pass 1 in the stacked column-march form (stack_item): stacks of four waves
// of R rows on plane blockIdx.y.
template <typename T, int ORDER, int R>
__global__ void __launch_bounds__(TPB) k_filter3d_plane_s(FilterArgs<T> a, size_t plane) {
  __shared__ StackLds<T, 4, R> L;
  const int t = xcd_tile(blockIdx.x, a.chunk);
  if (t >= a.ntiles) return;
  FilterArgs<T> b = a;
  b.src = a.src + (size_t)blockIdx.y * plane;
  b.dst = a.dst + (size_t)blockIdx.y * plane;
  stack_item<T, ORDER, 4, R>(b, t, L);
}

// Single pass: spatial::apply_filter's three sweeps (spatial:2210-2237,
// xi, then eta, then zeta) for one variable on a 64 x TY column tile
// through a chunk of planes.  Each plane of the tile plus its 4-point (i, j)
// halo goes through LDS for the xi and eta sweeps (the next plane's loads
// are in flight meanwhile); the thread keeps the xi-eta results of its TY/4
// points for the last 9 planes in registers, and when plane k + 4 has
// entered the window, sweeps zeta at plane k and blends with the unfiltered
// value of plane k, kept in a 5-plane register ring since its load.  A chunk
// reads its 4 halo planes on each side again (they are the neighbouring
// chunks' planes); nothing else is read twice and no intermediate goes to
// memory.  Each point takes the operations of the two-pass form in the
// same order, so the results are bit-identical to it.
template <typename T, int TY>
using PlaneRegs = FilterRegs<T, 1, TY>;

// LDS of the single pass: the plane's input and its xi sweep apart (the 2-D
// tile overlays them), so a plane takes two barriers: after the input is
// stored and after the xi rows are.  The next plane stores its input once
// every thread has passed the second barrier, i.e. finished its xi reads,
// and its xi rows once every thread has passed its first, i.e. finished
// this plane's eta reads.
template <typename T, int TY>
struct alignas(32) PlaneLds {
  T in[1][TY + 2 * FL_H][FL_TX + 2 * FL_H];
  T mid[1][TY + 2 * FL_H][FL_TX];
};

template <typename T, int ORDER, int TY, bool INNER>
__device__ __forceinline__ void plane_xieta(const FilterArgs<T> &a, int tile, PlaneLds<T, TY> &L,
                                            const PlaneRegs<T, TY> &R, T (&f)[TY / 4], T (&orig)[TY / 4]) {
  constexpr int NR = TY + 2 * FL_H, KR = NR / 4, NB = ORDER / 2;
  auto &in = L.in;
  auto &mid = L.mid;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj;
  const bool xtra = tx < 2 * FL_H;
  const int colb = xtra ? FL_TX + tx : tx;
  const int gia = i0 - FL_H + tx, gib = i0 - FL_H + colb;
  const bool cola = INNER || (gia >= 0 && gia < ni), colok_b = INNER || (gib >= 0 && gib < ni);
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int gj = j0 - FL_H + ty + 4 * k;
    const bool rowok = INNER || (gj >= 0 && gj < nj);
    in[0][ty + 4 * k][tx] = rowok && cola ? R.ra[0][k] : (T)0;
    in[0][ty + 4 * k][colb] = rowok && colok_b ? R.rb[0][k] : (T)0;
  }
  __syncthreads();
  constexpr int NG = FL_TX / 4;
  constexpr int NXI = NR * NG;
  constexpr int NS = (NXI + TPB - 1) / TPB;
  const int tid = ty * FL_TX + tx;
  Vec4<T> xo[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int it = tid + s * TPB;
    if (NXI % TPB != 0 && it >= NXI) break;
    const int g = it % NG, r = it / NG;
    T v[12];
#pragma unroll
    for (int h = 0; h < 3; h++) {
      const Vec4<T> x = *reinterpret_cast<const Vec4<T> *>(&in[0][r][4 * g + 4 * h]);
      v[4 * h] = x.x;
      v[4 * h + 1] = x.y;
      v[4 * h + 2] = x.z;
      v[4 * h + 3] = x.w;
    }
    const int gj = j0 - FL_H + r;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int gi = i0 + 4 * g + m;
      int code = CODE_INTERIOR;
      bool ok = true;
      if (!INNER) {
        ok = gj >= 0 && gj < nj && gi < ni;
        if (ok) code = default_code(gi, ni, NB);
      }
      xo[s][m] = ok ? filter_row<T, ORDER>(&v[m + FL_H], 1, code) : (T)0;
    }
    *reinterpret_cast<Vec4<T> *>(&mid[0][r][4 * g]) = xo[s];
  }
  __syncthreads();
  const int gi = i0 + tx;
  const int src_lane = tx < FL_TX - FL_H ? tx + FL_H : tx - (FL_TX - FL_H);
#pragma unroll
  for (int k = 0; k < TY / 4; k++) {
    const T va = __shfl(R.ra[0][k + 1], src_lane, FL_TX);
    const T vb = __shfl(R.rb[0][k + 1], src_lane, FL_TX);
    orig[k] = tx < FL_TX - FL_H ? va : vb;
    const int r = ty + 4 * k;
    const int gj = j0 + r;
    const int code = INNER ? CODE_INTERIOR : ((gj < nj && gi < ni) ? default_code(gj, nj, NB) : CODE_ZERO);
    f[k] = filter_row<T, ORDER>(&mid[0][r + FL_H][tx], FL_TX, code);
  }
}

template <typename T, int ORDER, int TY>
__global__ void __launch_bounds__(TPB) k_filter3d_fused(FilterArgs<T> a, int nk, int zchunk) {
  constexpr int NP = TY / 4, H = FL_H, NB = ORDER / 2;
  __shared__ PlaneLds<T, TY> L;
  const int t = xcd_tile(blockIdx.x, a.chunk);
  if (t >= a.ntiles) return;
  const int per_var = a.ntx * a.nty;
  const int v = t / per_var, tile = t - v * per_var;
  const size_t plane = (size_t)a.ni * a.nj;
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  const bool inner = i0 >= FL_H && i0 + FL_TX + FL_H <= a.ni && j0 >= FL_H && j0 + TY + FL_H <= a.nj;
  const int k0 = blockIdx.y * zchunk, k1 = min(k0 + zchunk, nk);
  FilterArgs<T> b = a;
  b.vars[0] = 0;
  b.vs = 0;
  const T *src = a.src + (size_t)v * a.vs;
  T *dst = a.dst + (size_t)v * a.vs;
  T w[NP][2 * H + 1], yo[NP][H + 1];
#pragma unroll
  for (int p = 0; p < NP; p++) {
#pragma unroll
    for (int q = 0; q <= 2 * H; q++) w[p][q] = (T)0;
#pragma unroll
    for (int q = 0; q <= H; q++) yo[p][q] = (T)0;
  }
  const int kfirst = max(k0 - H, 0), klast = min(k1 + H, nk) - 1;
  const int gi = i0 + threadIdx.x;
  // One plane kk: `cur` holds plane clamp(kk, kfirst, klast) and the loads
  // of plane clamp(kk + 1, ...) into `nxt` are issued first, whatever the
  // branch below takes, so that they stay in flight through this plane's
  // sweeps (a load issued on one path only makes the compiler wait for
  // every outstanding load before the sweeps).  The loop runs an even
  // number of planes, alternating the two register sets, so that no copy
  // between them waits for the loads behind this plane's stores.
  auto step = [&](int kk, const PlaneRegs<T, TY> &cur, PlaneRegs<T, TY> &nxt) {
    T f[NP], orig[NP];
    b.src = src + (size_t)min(max(kk + 1, kfirst), klast) * plane;
    filter_load<T, false, 1, TY>(b, tile, nxt);
    __builtin_amdgcn_sched_barrier(0);
    if (kk >= kfirst && kk <= klast) {
      if (inner) plane_xieta<T, ORDER, TY, true>(b, tile, L, cur, f, orig);
      else plane_xieta<T, ORDER, TY, false>(b, tile, L, cur, f, orig);
    } else {
#pragma unroll
      for (int p = 0; p < NP; p++) f[p] = orig[p] = (T)0;
    }
#pragma unroll
    for (int p = 0; p < NP; p++) {
#pragma unroll
      for (int q = 0; q < 2 * H; q++) w[p][q] = w[p][q + 1];
      w[p][2 * H] = f[p];
#pragma unroll
      for (int q = 0; q < H; q++) yo[p][q] = yo[p][q + 1];
      yo[p][H] = orig[p];
    }
    const int k = kk - H;
    if (k >= k0 && k < k1) {
      const int code = default_code(k, nk, NB);
#pragma unroll
      for (int p = 0; p < NP; p++) {
        const int gj = j0 + threadIdx.y + 4 * p;
        if (!inner && (gj >= a.nj || gi >= a.ni)) continue;
        const T fz = filter_row<T, ORDER>(&w[p][H], 1, code);
        dst[(size_t)k * plane + (size_t)gj * a.ni + gi] = filter_blend(a.rem, yo[p][0], a.blend, fz);
      }
    }
  };
  PlaneRegs<T, TY> A, B;
  b.src = src + (size_t)kfirst * plane;
  filter_load<T, false, 1, TY>(b, tile, A);
  // planes k0 - H .. k1 + H - 1, plus one past them when that count is odd
  const int kend = k0 - H + ((k1 - k0 + 2 * H + 1) & ~1);
  for (int kk = k0 - H; kk < kend; kk += 2) {
    step(kk, A, B);
    step(kk + 1, B, A);
  }
}

// Pass 2: the zeta sweep of tmp along each (i, j) column through a chunk of
// planes (9-plane register window), blend with y, into dst.
template <typename T, int ORDER>
__global__ void __launch_bounds__(TPB) k_filter3d_z(const T *__restrict__ tmp, const T *__restrict__ y,
                                                    T *__restrict__ dst, int ni, int nj, int nk, T rem, T blend) {
  constexpr int H = FL_H, NB = ORDER / 2;
  const size_t plane = (size_t)ni * nj, n = plane * nk;
  const size_t col = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (col >= plane) return;
  const int k0 = blockIdx.y * Z_CHUNK, k1 = min(k0 + Z_CHUNK, nk);
  T w[5][2 * H + 1];
#pragma unroll
  for (int q = 0; q < 2 * H; q++) {
    const int kk = k0 - H + q;
    const bool ok = kk >= 0 && kk < nk;
#pragma unroll
    for (int v = 0; v < 5; v++) w[v][q] = ok ? tmp[v * n + (size_t)kk * plane + col] : (T)0;
  }
  // four planes per pass: their window planes and unfiltered values are
  // requested together (clamped plane index, zeros selected afterwards),
  // one memory latency per four planes
  auto plane_out = [&](int k, const T (&wn)[5], const T (&yk)[5]) {
#pragma unroll
    for (int v = 0; v < 5; v++) w[v][2 * H] = wn[v];
    const int code = default_code(k, nk, NB);
    const size_t p = (size_t)k * plane + col;
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const T f = filter_row<T, ORDER>(&w[v][H], 1, code);
      dst[v * n + p] = filter_blend(rem, yk[v], blend, f);
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
      for (int q = 0; q < 2 * H; q++) w[v][q] = w[v][q + 1];
  };
  int k = k0;
  for (; k + 4 <= k1; k += 4) {
    T wn[4][5], yk[4][5];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = k + u + H;
      const size_t pc = (size_t)min(kk, nk - 1) * plane + col, pk = (size_t)(k + u) * plane + col;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const T t = tmp[v * n + pc];
        wn[u][v] = kk < nk ? t : (T)0;
        yk[u][v] = y[v * n + pk];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) plane_out(k + u, wn[u], yk[u]);
  }
  for (; k < k1; k++) {
    const int kk = k + H;
    T wn[5], yk[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      wn[v] = kk < nk ? tmp[v * n + (size_t)kk * plane + col] : (T)0;
      yk[v] = y[v * n + (size_t)k * plane + col];
    }
    plane_out(k, wn, yk);
  }
}

}  // namespace

// ------------------------------------------------------------- launchers

hipError_t metrics3d(hipStream_t s, Block &b) {
  double *xx = nullptr;
  const size_t n = b.n;
  hipError_t e = hipMalloc((void **)&xx, sizeof(double) * 9 * n);
  if (e != hipSuccess) return e;
  for (int i = 0; i < 3; i++)
    for (int l = 0; l < 3; l++)
      hipLaunchKernelGGL(k_op_dir3, dim3(nb3(n)), dim3(TPB), 0, s, b.xyz + (size_t)i * n, xx + (size_t)(3 * i + l) * n,
                         b.ni, b.nj, b.nk, l);
  hipLaunchKernelGGL(k_metrics3d, dim3(nb3(n)), dim3(TPB), 0, s, xx, b.met, b.jac, n);
  e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(xx);
  return e != hipSuccess ? e : e2;
}

hipError_t deriv_xyz3(hipStream_t s, const Block &b, const double *f, double *out, int d) {
  hipLaunchKernelGGL(k_deriv_xyz3, dim3(nb3(b.n)), dim3(TPB), 0, s, f, out, b.met, b.jac, b.ni, b.nj, b.nk, d);
  return hipGetLastError();
}

template <typename T>
static hipError_t stage3d_t(hipStream_t s, lee_handle *h, Block &b, int stage, const StageScalars &sc, T *const *Y,
                            T *rho) {
  const int last = h->cfg.rk_stages - 1;
  const int kind = stage <= 0 ? (stage == 0 ? STAGE_FIRST : STAGE_MID) : (stage == last ? STAGE_LAST : STAGE_MID);
  Stage3Args<T> a{};
  a.ni = b.ni;
  a.nj = b.nj;
  a.nk = b.nk;
  a.ntx = (b.ni + ST_TX - 1) / ST_TX;
  a.nty = (b.nj + S3_TY - 1) / S3_TY;
  a.nzc = (b.nk + Z_CHUNK - 1) / Z_CHUNK;
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
  a.mc = sizeof(T) == 8 ? (const T *)b.mc64 : (const T *)b.mc32;
  a.rho_in = rho;
  a.rho_out = rho;
  a.u0 = b.umean[IUX];
  a.v0 = b.umean[IUY];
  a.w0 = b.umean[IUZ];
  a.gamma = h->cfg.gamma;
  a.gp0 = h->cfg.gamma * b.umean[IP];
  a.rhob0 = std::pow(h->cfg.gamma * b.umean[IP], 1.0 / h->cfg.gamma) * std::exp(-b.umean[IS]);
  a.nz = 0;
  for (const auto &z : b.zones) {
    if (a.nz >= MAX_ZONES) break;
    Zone3<T> &q = a.z[a.nz++];
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
  a.si0 = b.src.i0;
  a.si1 = b.src.i1;
  a.sj0 = b.src.j0;
  a.sj1 = b.src.j1;
  a.sk0 = b.src.k0;
  a.sk1 = b.src.k1;
  a.g = b.src.g;
  a.s1 = sc.s1;
  a.s2 = sc.s2;
  a.h = (T)sc.h;
  a.b = (T)sc.b;
  a.a = (T)sc.a;
  const dim3 grid((a.ntx * a.nty + 7) / 8 * 8, a.nzc), block(ST_TX, S3_TY);
  const bool mf = !b.uniform_mean;
#define LEE_S3(MF)                                                                                    \
  do {                                                                                                \
    if (kind == STAGE_FIRST) hipLaunchKernelGGL((k_stage3d<T, MF, STAGE_FIRST>), grid, block, 0, s, a); \
    else if (kind == STAGE_MID) hipLaunchKernelGGL((k_stage3d<T, MF, STAGE_MID>), grid, block, 0, s, a); \
    else hipLaunchKernelGGL((k_stage3d<T, MF, STAGE_LAST>), grid, block, 0, s, a);                     \
  } while (0)
  if (mf) LEE_S3(true);
  else LEE_S3(false);
#undef LEE_S3
  return hipGetLastError();
}

hipError_t fused_stage3d(hipStream_t s, lee_handle *h, Block &b, int stage, const StageScalars &sc) {
  if (h->cfg.precision == LEE_FP32) return stage3d_t<float>(s, h, b, stage, sc, b.Y32, b.rho32);
  return stage3d_t<double>(s, h, b, stage, sc, b.Y64, b.rho64);
}

template <typename T>
static hipError_t filter3d_t(hipStream_t s, lee_handle *h, Block &b, T *const *Y, int src, int dst, int tmp) {
  const int TYF = sizeof(T) == 8 ? 4 : 8;
  FilterArgs<T> a{};
  a.ni = b.ni;
  a.nj = b.nj;
  a.vs = (size_t)b.n;
  a.ntx = (b.ni + FL_TX - 1) / FL_TX;
  a.nty = (b.nj + TYF - 1) / TYF;
  a.ntiles = a.ntx * a.nty;
  a.chunk = (a.ntiles + 7) / 8;
  a.src = Y[src];
  a.dst = Y[tmp];
  a.blend = (T)1;
  a.rem = (T)0;
  for (int v = 0; v < 5; v++) a.vars[v] = v;
  const size_t plane = (size_t)b.ni * b.nj;
  const bool o8 = h->cfg.filter_order == 8;
  if (b.tune_mc == 1) {  // single pass, 64 x tune_ml column tiles, one variable each
    const int TY = b.tune_ml;
    a.dst = Y[dst];
    a.blend = (T)h->cfg.filter_blend;
    a.rem = (T)(1.0 - h->cfg.filter_blend);
    a.nty = (b.nj + TY - 1) / TY;
    a.ntiles = 5 * a.ntx * a.nty;
    a.chunk = (a.ntiles + 7) / 8;
    const int zc = FUSED_Z_CHUNK;
    const dim3 grid(a.chunk * 8, (b.nk + zc - 1) / zc), block(64, 4);
#define LEE_F3F(OR)                                                                                      \
  do {                                                                                                   \
    if (TY == 8) hipLaunchKernelGGL((k_filter3d_fused<T, OR, 8>), grid, block, 0, s, a, b.nk, zc);     \
    else hipLaunchKernelGGL((k_filter3d_fused<T, OR, 16>), grid, block, 0, s, a, b.nk, zc);            \
  } while (0)
    if (o8) LEE_F3F(8);
    else LEE_F3F(4);
#undef LEE_F3F
    return hipGetLastError();
  }
  if (b.tune_mc == 4) {  // stacks of four waves of tune_ml rows
    const int R = b.tune_ml;
    a.ntx = (b.ni + March::STRIDE - 1) / March::STRIDE;
    a.nty = (b.nj + 4 * R - 1) / (4 * R);
    a.ntiles = 5 * a.ntx * a.nty;
    a.chunk = (a.ntiles + 7) / 8;
    const dim3 grid(a.chunk * 8, b.nk), block(64, 4);
#define LEE_F3S(OR)                                                                                   \
  do {                                                                                                \
    if (R == 4) hipLaunchKernelGGL((k_filter3d_plane_s<T, OR, 4>), grid, block, 0, s, a, plane);      \
    else if (R == 8) hipLaunchKernelGGL((k_filter3d_plane_s<T, OR, 8>), grid, block, 0, s, a, plane); \
    else hipLaunchKernelGGL((k_filter3d_plane_s<T, OR, 16>), grid, block, 0, s, a, plane);            \
  } while (0)
    if (o8) LEE_F3S(8);
    else LEE_F3S(4);
#undef LEE_F3S
  } else {
    const dim3 grid(a.chunk * 8, b.nk), block(64, 4);
    if (o8) hipLaunchKernelGGL((k_filter3d_plane<T, 8, TYF>), grid, block, 0, s, a, plane);
    else hipLaunchKernelGGL((k_filter3d_plane<T, 4, TYF>), grid, block, 0, s, a, plane);
  }
  const dim3 gz(nb3(plane), (b.nk + Z_CHUNK - 1) / Z_CHUNK);
  const T blend = (T)h->cfg.filter_blend, rem = (T)(1.0 - h->cfg.filter_blend);
  if (o8) hipLaunchKernelGGL((k_filter3d_z<T, 8>), gz, dim3(TPB), 0, s, Y[tmp], Y[src], Y[dst], b.ni, b.nj, b.nk, rem, blend);
  else hipLaunchKernelGGL((k_filter3d_z<T, 4>), gz, dim3(TPB), 0, s, Y[tmp], Y[src], Y[dst], b.ni, b.nj, b.nk, rem, blend);
  return hipGetLastError();
}

hipError_t fused_filter3d(hipStream_t s, lee_handle *h, Block &b, int src, int dst, int tmp) {
  if (h->cfg.precision == LEE_FP32) return filter3d_t<float>(s, h, b, b.Y32, src, dst, tmp);
  return filter3d_t<double>(s, h, b, b.Y64, src, dst, tmp);
}

}  // namespace lee
