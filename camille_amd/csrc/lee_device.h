// Device-side definitions shared by the forward and adjoint kernels of the
// LEE time advance: tile geometry, operator rows as compile-time
// constants, and the per-launch operand structs.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lee_internal.h"

namespace lee {

namespace {

constexpr int TPB = 256;

// Read-only tables that kernels index at run time (the batched ensemble's
// per-solve operands) are declared in the constant address space: loads
// from it are scalar and the pointers they yield address global memory.
#define LEE_CONST_AS __attribute__((address_space(4)))

__device__ __forceinline__ int default_code(int c, int len, int nb) {
  if (c < nb) return CODE_LEFT + c;
  if (c > len - 1 - nb) return CODE_LEFT + nb + (len - 1 - c);
  return CODE_INTERIOR;
}


constexpr int ST_TX = 64, ST_H = 2;   // stage tile width and FD halo
constexpr int FL_TX = 64, FL_H = 4;   // filter tile width and halo (height: autotuned)

// The optimized 11-point schemes (FDO11P derivative, SFO11P filter) in the
// fused kernels: an OpTable in device memory, each closure row with its own
// length.  Their tiles take a halo of 6: the interior rows reach 5 points,
// FDO11P's closure rows 0 and 1 (anchored at the boundary, 7 points) reach
// 6 from the point, which a closure next to a hole inside a tile needs.
constexpr int WIDE_H = 6;

// Halo and closure-row count of the stage's derivative operator: OA4 (W 4)
// or FDO11P (W 11).
template <int W> struct StageOp;
template <> struct StageOp<4> { static constexpr int H = ST_H, NB = 2; };
template <> struct StageOp<11> { static constexpr int H = WIDE_H, NB = 5; };

// OA4 finite-difference rows of StandardCentral::initialize (spatial:64-105)
// as compile-time constants: interior {1/12, -8/12, 0, 8/12, -1/12}; left
// rows 0 {-3/2, 2, -1/2} (anchored at the point) and 1 {-1/2, 0, 1/2}
// (anchored one point left); right rows mirrored with the sign flipped.
// Whether the box [i0, i1) x [j0, j1) lies inside an ni x nj block and holds
// no point flagged in the integral image S ((ni + 1) x (nj + 1), S[j][i] =
// flagged points with j' < j and i' < i).  Uniform per workgroup.
__device__ __forceinline__ bool box_clean(const int *S, int ni, int nj, int i0, int i1, int j0, int j1) {
  if (i0 < 0 || j0 < 0 || i1 > ni || j1 > nj) return false;
  const int w = ni + 1;
  return S[j1 * w + i1] - S[j0 * w + i1] - S[j1 * w + i0] + S[j0 * w + i0] == 0;
}

// Explicit fused multiply-add: the fused kernels spell out every fusion so
// that their instances (interior tiles, edge tiles, batched) round alike.
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One row of a WideOp at the point s (stride st along the sweep): the
// interior row, zero (hole after hole), a left closure row anchored k
// points to the left, or a right closure row mirrored with sign_right.
template <typename T>
__device__ __forceinline__ T wide_row(const T *s, int st, int code, const WideOp *op) {
#pragma clang fp contract(off)
  if (code == CODE_INTERIOR) {
    T acc = (T)op->st[0] * s[-5 * st];
#pragma unroll
    for (int k = 1; k < 11; k++) acc = fmx((T)op->st[k], s[(k - 5) * st], acc);
    return acc;
  }
  if (code == CODE_ZERO) return (T)0;
  int k = code - CODE_LEFT;
  T acc = 0;
  if (k < op->nb) {
    for (int q = 0; q < op->len[k]; q++) acc = fmx((T)op->bnd[k][q], s[(q - k) * st], acc);
    return acc;
  }
  k -= op->nb;
  for (int q = 0; q < op->len[k]; q++) acc = fmx((T)op->bnd[k][q], s[(k - q) * st], acc);
  return (T)op->sign_right * acc;
}

template <typename T>
__device__ __forceinline__ T fd_row(const T *s, int st, int code) {
#pragma clang fp contract(off)
  const T c1 = (T)(8.0 / 12.0), c2 = (T)(-1.0 / 12.0);
  switch (code) {
    case CODE_INTERIOR: {
      T acc = -c2 * s[-2 * st];
      acc = fmx(-c1, s[-st], acc);
      acc = fmx(c1, s[st], acc);
      acc = fmx(c2, s[2 * st], acc);
      return acc;
    }
    case CODE_ZERO:
      return (T)0;
    case CODE_LEFT + 0: {
      T acc = (T)(-3.0 / 2.0) * s[0];
      acc = fmx((T)(4.0 / 2.0), s[st], acc);
      acc = fmx((T)(-1.0 / 2.0), s[2 * st], acc);
      return acc;
    }
    case CODE_LEFT + 1: {
      T acc = (T)(-1.0 / 2.0) * s[-st];
      acc = fmx((T)(1.0 / 2.0), s[st], acc);
      return acc;
    }
    case CODE_LEFT + 2: {
      T acc = (T)(3.0 / 2.0) * s[0];
      acc = fmx((T)(-4.0 / 2.0), s[-st], acc);
      acc = fmx((T)(1.0 / 2.0), s[-2 * st], acc);
      return acc;
    }
    default: {
      T acc = (T)(1.0 / 2.0) * s[st];
      acc = fmx((T)(-1.0 / 2.0), s[-st], acc);
      return acc;
    }
  }
}

template <typename T>
struct ZoneT {
  int i0, i1, j0, j1, nzi;
  const T *sig;
  int harm;            // sponge_freund_harmonicwave: s' relaxes to he * hg (hg null: he)
  const double *hg;
};

template <typename T>
struct StageArgs {
  int ni, nj, ntx, nty, ntiles, chunk;
  const T *yin;        // stage input (5n, var-major)
  const T *yn;         // y_n
  T *acc;              // y_next accumulator (read-modify-write)
  T *yout;             // next stage input
  const double *met, *jac;
  const int8_t *iblank, *code;  // holes only
  const int *clean;    // holes only: integral image of points off the interior stencil
  const double *mean, *mgrad, *rhob;   // non-uniform mean only
  const T *mc;         // non-uniform mean: MC_COUNT coefficient fields
  const T *rho_in;     // stale rho' (first stage, non-uniform mean)
  T *rho_out;          // rho' of the new y_{n+1} (last stage, non-uniform mean)
  double u0, v0, gp0, rhob0, gamma, gi;
  int axi;             // axisymmetric pressure term (non-uniform mean path)
  int nz;
  ZoneT<T> z[MAX_ZONES];
  int si0, si1, sj0, sj1;
  const double *g;
  double s1, s2;
  double he;           // harmonic-wave entropy factor at the stage time
  T h, b, a;
  const WideOp *fdw;   // FDO11P rows (the 11-point stage instances only)
};

// Maps the launch index to a tile so that each XCD (dispatch is round-robin
// over the 8 XCDs) walks a contiguous band of tile rows and keeps the halo
// rows of vertically adjacent tiles in its own L2.
__device__ __forceinline__ int xcd_tile(int bid, int chunk) { return (bid & 7) * chunk + (bid >> 3); }

// One RK stage on a 64 x TY tile.  Every global operand of the thread (its
// share of the LDS tile, the FP64 metrics and Jacobian, the accumulator and
// y_n of its R points) is requested up front, so a launch that is a single
// wave of tiles on the chip waits on HBM once instead of once per row.
// R rows per thread (tile height 4 R) is chosen per block by the autotuner.
template <typename T>
struct StepT {       // per-stage scalars: dt, b_i, a_{i+1,i}, source time factors
  T h, b, a;
  double s1, s2;
};

template <typename T, int R>
using StageLds = T[4][4 * R + 2 * ST_H][ST_TX + 2 * ST_H];
template <typename T, int R, int W>
using StageLdsW = T[4][4 * R + 2 * StageOp<W>::H][ST_TX + 2 * StageOp<W>::H];


template <typename T>
struct FilterArgs {
  int ni, nj, ntx, nty, ntiles, chunk;
  const T *src;      // unfiltered y (5n)
  T *dst;            // filtered y (5n)
  const int8_t *iblank, *code;  // holes only (filter codes, xi then eta)
  const int *clean;  // holes only: integral image of points off the interior filter
  int vars[5];       // variables of this launch
  size_t vs;         // variable stride (0: ni nj)
  T rem, blend;
  const WideOp *flw; // SFO11P rows (the 11-point filter instances only)
};

// Filter rows with the reference's coefficients as compile-time constants
// (StandardCentralFilter::initialize, spatial:113-224; the standard filter
// always runs at strength 1, input.cpp:50-53, so the tables are fixed).
// ORDER 8: interior 1 - 35/128 at the centre and 7/32, -7/64, 1/32, -1/256
// off it; rows 0, 1 identity, row 2 Lele's 4th-order and row 3 his
// 6th-order centred filter.  ORDER 4: Lele's 4th-order filter, rows 0, 1
// identity.  Right rows are the mirror images (sign +1).  `s` points at the
// point, `st` is the stride along the sweep.
template <typename T, int ORDER>
__device__ __forceinline__ T filter_row(const T *s, int st, int code) {
#pragma clang fp contract(off)
  // every fusion is spelled out, so that the tile forms and the column-march
  // form round each point alike
  if (ORDER == 8) {
    if (code == CODE_INTERIOR) {
      const T c0 = (T)(93.0 / 128.0), c1 = (T)(7.0 / 32.0), c2 = (T)(-7.0 / 64.0), c3 = (T)(1.0 / 32.0),
              c4 = (T)(-1.0 / 256.0);
      T acc = c4 * s[-4 * st];
      acc = fmx(c3, s[-3 * st], acc);
      acc = fmx(c2, s[-2 * st], acc);
      acc = fmx(c1, s[-st], acc);
      acc = fmx(c0, s[0], acc);
      acc = fmx(c1, s[st], acc);
      acc = fmx(c2, s[2 * st], acc);
      acc = fmx(c3, s[3 * st], acc);
      acc = fmx(c4, s[4 * st], acc);
      return acc;
    }
    if (code == CODE_ZERO) return (T)0;
    const int k = code - CODE_LEFT;
    const int row = k < 4 ? k : k - 4;
    if (row <= 1) return s[0];
    if (row == 2) {
      const T a = (T)((-1.0 / 8.0) / 2.0), b = (T)((1.0 / 2.0) / 2.0), c = (T)(5.0 / 8.0);
      T acc = a * s[-2 * st];
      acc = fmx(b, s[-st], acc);
      acc = fmx(c, s[0], acc);
      acc = fmx(b, s[st], acc);
      acc = fmx(a, s[2 * st], acc);
      return acc;
    }
    const T a = (T)((1.0 / 32.0) / 2.0), b = (T)((-3.0 / 16.0) / 2.0), c = (T)((15.0 / 32.0) / 2.0),
            d = (T)(11.0 / 16.0);
    T acc = a * s[-3 * st];
    acc = fmx(b, s[-2 * st], acc);
    acc = fmx(c, s[-st], acc);
    acc = fmx(d, s[0], acc);
    acc = fmx(c, s[st], acc);
    acc = fmx(b, s[2 * st], acc);
    acc = fmx(a, s[3 * st], acc);
    return acc;
  } else {
    if (code == CODE_INTERIOR) {
      const T a = (T)((-1.0 / 8.0) / 2.0), b = (T)((1.0 / 2.0) / 2.0), c = (T)(5.0 / 8.0);
      T acc = a * s[-2 * st];
      acc = fmx(b, s[-st], acc);
      acc = fmx(c, s[0], acc);
      acc = fmx(b, s[st], acc);
      acc = fmx(a, s[2 * st], acc);
      return acc;
    }
    if (code == CODE_ZERO) return (T)0;
    return s[0];
  }
}

// The filter's blend of the unfiltered value u with the filtered f
// (rem = 1 - blend), fused as one FMA.
template <typename T>
__device__ __forceinline__ T filter_blend(T rem, T u, T blend, T f) {
#pragma clang fp contract(off)
  return fmx(rem, u, blend * f);
}

template <typename T, int NVAR, int TY, int H = FL_H>
union alignas(32) FilterLds {  // the xi-sweep overwrites its input
  T in[NVAR][TY + 2 * H][FL_TX + 2 * H];
  T mid[NVAR][TY + 2 * H][FL_TX];
};

// Halo of the filter tiles: 4 for the standard filters, WIDE_H for SFO11P.
template <int ORDER>
constexpr int filter_halo() { return ORDER == 11 ? WIDE_H : FL_H; }

// A filter row of ORDER 4 or 8 (compile-time rows) or of SFO11P (a.flw).
template <typename T, int ORDER, typename AT>
__device__ __forceinline__ T filter_row_of(const AT &a, const T *s, int st, int code) {
  if constexpr (ORDER == 11) return wide_row<T>(s, st, code, a.flw);
  else return filter_row<T, ORDER>(s, st, code);
}

// Four consecutive LDS values as one vector access (rows of the filter
// tiles are padded to multiples of four values).
template <typename T>
using Vec4 = T __attribute__((ext_vector_type(4)));

// spatial::apply_filter (spatial:2191-2246) on a 64 x TY tile for NVAR
// variables at once: every global operand is requested up front and the
// tile plus its 4-point halo goes to LDS; the xi sweep of every region row
// (4 consecutive columns per item from three 4-vector LDS reads) lands in
// registers and then overwrites the tile in place; the eta sweep, the blend
// with the unfiltered values (kept in registers since the load) and
// temporal::iblank_correction follow.  Tiles whose halo stays inside the
// block skip all edge logic.
// Global loads of one filter tile (tile plus its 4-point halo) into the
// thread's registers: region rows ty + 4k, columns tx and, for tx < 8,
// 64 + tx.
template <typename T, int NVAR, int TY, int H = FL_H>
struct FilterRegs {
  T ra[NVAR][(TY + 2 * H) / 4], rb[NVAR][(TY + 2 * H) / 4];
};

template <typename T, bool HOLES, int NVAR, int TY, int H = FL_H, typename AT>
__device__ __forceinline__ void filter_load(const AT &a, int tile, FilterRegs<T, NVAR, TY, H> &R) {
  constexpr int NR = TY + 2 * H, KR = NR / 4;
  static_assert(NR == 4 * KR && H % 2 == 0, "tile rows and the halo come in fours");
  auto &ra = R.ra;
  auto &rb = R.rb;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const size_t vs = a.vs ? a.vs : (size_t)n;
  const bool xtra = tx < 2 * H;
  // one straight-line sequence of loads for every tile, from clamped
  // addresses (filter_compute puts zeros outside the block): the wait
  // counters then let a caller keep the next tile's loads in flight while it
  // filters the current one.  Lanes without a right-halo column repeat
  // their own column.
  const int gia = i0 - H + tx, gib = xtra ? i0 - H + FL_TX + tx : gia;
  const int cia = min(max(gia, 0), ni - 1), cib = min(max(gib, 0), ni - 1);
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int gj = j0 - H + ty + 4 * k;
    const int cj = min(max(gj, 0), nj - 1);
    const int pa = cj * ni + cia, pb = cj * ni + cib;
#pragma unroll
    for (int q = 0; q < NVAR; q++) {
      const T *src = a.src + a.vars[q] * vs;
      ra[q][k] = src[pa];
      rb[q][k] = src[pb];
    }
  }
}

// The filter of one tile from the registers of filter_load (see filter_tile).
// INNER: the tile and its halo lie inside a block without holes, so every
// point takes the interior stencil and no bounds are tested.
template <typename T, bool HOLES, int ORDER, int NVAR, int TY, bool INNER, typename AT>
__device__ __forceinline__ void filter_compute_k(const AT &a, int tile,
                                                 FilterLds<T, NVAR, TY, filter_halo<ORDER>()> &L,
                                                 const FilterRegs<T, NVAR, TY, filter_halo<ORDER>()> &R) {
  constexpr int H = filter_halo<ORDER>();
  constexpr int NR = TY + 2 * H, KR = NR / 4;
  constexpr int NB = ORDER == 11 ? 5 : ORDER / 2;
  constexpr int NV4 = (4 + 2 * H) / 4;  // 4-vectors of LDS behind 4 xi outputs
  auto &in = L.in;
  auto &mid = L.mid;
  const auto &ra = R.ra;
  const auto &rb = R.rb;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  const int ni = a.ni, nj = a.nj, n = ni * nj;
  const size_t vs = a.vs ? a.vs : (size_t)n;
  const bool xtra = tx < 2 * H;
  constexpr bool inner = INNER;
  // rb of a lane without a right-halo column holds its own column again
  // (filter_load): writing it back there keeps the stores unconditional
  const int colb = xtra ? FL_TX + tx : tx;
  const int gia = i0 - H + tx, gib = i0 - H + colb;
  const bool cola = inner || (gia >= 0 && gia < ni), colok_b = inner || (gib >= 0 && gib < ni);
#pragma unroll
  for (int k = 0; k < KR; k++) {
    const int gj = j0 - H + ty + 4 * k;
    const bool rowok = inner || (gj >= 0 && gj < nj);
    const bool oka = rowok && cola, okb = rowok && colok_b;
#pragma unroll
    for (int q = 0; q < NVAR; q++) {
      in[q][ty + 4 * k][tx] = oka ? ra[q][k] : (T)0;
      in[q][ty + 4 * k][colb] = okb ? rb[q][k] : (T)0;
    }
  }
  __syncthreads();
  // xi-sweep of every region row into registers: each item is 4 consecutive
  // columns of one row and variable, from 12 values read as three 4-vectors
  constexpr int NG = FL_TX / 4;
  constexpr int NXI = NVAR * NR * NG;
  constexpr int NS = (NXI + TPB - 1) / TPB;
  const int tid = ty * FL_TX + tx;
  Vec4<T> xo[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int it = tid + s * TPB;
    if (NXI % TPB != 0 && it >= NXI) break;
    const int g = it % NG, r = (it / NG) % NR, q = it / (NG * NR);
    T v[4 * NV4];
#pragma unroll
    for (int h = 0; h < NV4; h++) {
      const Vec4<T> x = *reinterpret_cast<const Vec4<T> *>(&in[q][r][4 * g + 4 * h]);
      v[4 * h] = x.x;
      v[4 * h + 1] = x.y;
      v[4 * h + 2] = x.z;
      v[4 * h + 3] = x.w;
    }
    const int gj = j0 - H + r;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const int gi = i0 + 4 * g + m;
      int code = CODE_INTERIOR;
      bool ok = true;
      if (!inner) {
        ok = gj >= 0 && gj < nj && gi < ni;
        if (ok) code = HOLES ? (int)a.code[gj * ni + gi] : default_code(gi, ni, NB);
      }
      xo[s][m] = ok ? filter_row_of<T, ORDER>(a, &v[m + H], 1, code) : (T)0;
    }
  }
  __syncthreads();  // every read of the input tile is done: mid overwrites it
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int it = tid + s * TPB;
    if (NXI % TPB != 0 && it >= NXI) break;
    const int g = it % NG, r = (it / NG) % NR, q = it / (NG * NR);
    *reinterpret_cast<Vec4<T> *>(&mid[q][r][4 * g]) = xo[s];
  }
  __syncthreads();
  // eta-sweep, blend with the unfiltered values kept in registers since the
  // load, and hole reset.  Thread (tx, ty) loaded region row ty + 4k, i.e.
  // tile row ty + 4(k - 1), at region column tx (block column i0 - 4 + tx)
  // in ra[.][k] and column 64 + tx in rb[.][k]: the value at block column
  // i0 + tx comes from lane tx + 4, or for tx >= 60 from rb of lane tx - 60.
  const int gi = i0 + tx;
  const int src_lane = tx < FL_TX - H ? tx + H : tx - (FL_TX - H);
#pragma unroll
  for (int k = 0; k < TY / 4; k++) {
    T orig[NVAR];
    if constexpr (H % 4 == 0) {  // the loaded row ty + 4 (k + H/4) is tile row ty + 4k
#pragma unroll
      for (int q = 0; q < NVAR; q++) {
        const T va = __shfl(ra[q][k + H / 4], src_lane, FL_TX);
        const T vb = __shfl(rb[q][k + H / 4], src_lane, FL_TX);
        orig[q] = tx < FL_TX - H ? va : vb;
      }
    }
    const int r = ty + 4 * k;
    const int gj = j0 + r;
    if (!inner && (gj >= nj || gi >= ni)) continue;
    const int p = gj * ni + gi;
    if constexpr (H % 4 != 0) {  // the halo of 6 shifts the loaded rows: the value again
#pragma unroll
      for (int q = 0; q < NVAR; q++) orig[q] = a.src[a.vars[q] * vs + p];
    }
    const int code = inner ? CODE_INTERIOR : (HOLES ? (int)a.code[n + p] : default_code(gj, nj, NB));
    const bool hole = HOLES && a.iblank[p] == BLANKED;
#pragma unroll
    for (int q = 0; q < NVAR; q++) {
      const T f = filter_row_of<T, ORDER>(a, &mid[q][r + H][tx], FL_TX, code);
      T out = filter_blend(a.rem, orig[q], a.blend, f);
      if (hole) out = 0;
      a.dst[a.vars[q] * vs + p] = out;
    }
  }
}

template <typename T, bool HOLES, int ORDER, int NVAR, int TY, typename AT>
__device__ __forceinline__ void filter_compute(const AT &a, int tile,
                                               FilterLds<T, NVAR, TY, filter_halo<ORDER>()> &L,
                                               const FilterRegs<T, NVAR, TY, filter_halo<ORDER>()> &R) {
  constexpr int H = filter_halo<ORDER>();
  const int i0 = (tile % a.ntx) * FL_TX, j0 = (tile / a.ntx) * TY;
  if constexpr (!HOLES) {
    if (i0 >= H && i0 + FL_TX + H <= a.ni && j0 >= H && j0 + TY + H <= a.nj) {
      filter_compute_k<T, HOLES, ORDER, NVAR, TY, true>(a, tile, L, R);
      return;
    }
  } else {
    // a tile of a block with holes whose points and halo rows all take the
    // interior filter (no hole, no closure row) runs the interior instance
    if (a.clean && box_clean(a.clean, a.ni, a.nj, i0, i0 + FL_TX, j0 - H, j0 + TY + H)) {
      filter_compute_k<T, false, ORDER, NVAR, TY, true>(a, tile, L, R);
      return;
    }
  }
  filter_compute_k<T, HOLES, ORDER, NVAR, TY, false>(a, tile, L, R);
}


template <typename T, bool HOLES, int ORDER, int NVAR, int TY, typename AT>
__device__ __forceinline__ void filter_tile(const AT &a, int tile,
                                            FilterLds<T, NVAR, TY, filter_halo<ORDER>()> &L) {
  FilterRegs<T, NVAR, TY, filter_halo<ORDER>()> R;
  filter_load<T, HOLES, NVAR, TY, filter_halo<ORDER>()>(a, tile, R);
  filter_compute<T, HOLES, ORDER, NVAR, TY>(a, tile, L, R);
}

// ---------------------------------------------------------------------------
// Column-march form of the filter (blocks without holes).  One wave filters
// one variable on a strip of 64 columns over L rows.  Lane l holds grid
// column i_s + l; the outer 4 lanes on each side are halo, so strips advance
// by 56 columns and overlap by 8.  The wave requests its L + 8 raw rows at
// once, then per row takes the xi neighbours from the adjacent lanes (DPP
// wave shifts: no LDS, no barrier), keeps the xi-filtered rows in registers
// and runs the eta sweep and the blend as soon as the nine rows of an output
// row are there.  Each point gets the operations of filter_tile in the same
// order, so both forms give the same bits.

// Whole-wave lane shifts by one lane (DPP wave_shr:1, control 0x138, and
// wave_shl:1, 0x130): UP gives lane l the value of lane l - 1, !UP that of
// lane l + 1; the lane without a source gets 0.
template <bool UP>
__device__ __forceinline__ int wave_shift1(int x) {
  return __builtin_amdgcn_update_dpp(0, x, UP ? 0x138 : 0x130, 0xF, 0xF, true);
}
template <bool UP>
__device__ __forceinline__ float wave_shift1(float x) {
  return __builtin_bit_cast(float, wave_shift1<UP>(__builtin_bit_cast(int, x)));
}
template <bool UP>
__device__ __forceinline__ double wave_shift1(double x) {
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = wave_shift1<UP>((int)(v & 0xffffffffll)), hi = wave_shift1<UP>((int)(v >> 32));
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Strip geometry of the column-march forms: one grid column per lane, the
// outer 4 lanes on each side are halo.
struct March {
  static constexpr int HL = FL_H;               // halo lanes on each side
  static constexpr int STRIDE = 64 - 2 * HL;    // output columns per strip
};

// One wave's share: variable slot q, strip s, row chunk c.  INNER: every
// load of the wave lies inside the block, so every output point takes the
// interior rows and no bounds are tested.
template <typename T, int ORDER, int L, bool INNER>
__device__ __forceinline__ void march_wave(const FilterArgs<T> &a, int q, int s, int c) {
  constexpr int NB = ORDER / 2, HL = March::HL, NR = L + 2 * FL_H;
  const int lane = threadIdx.x;
  const int ni = a.ni, nj = a.nj;
  const size_t vs = a.vs ? a.vs : (size_t)ni * nj;
  const T *src = a.src + a.vars[q] * vs;
  T *dst = a.dst + a.vars[q] * vs;
  const int gi = s * March::STRIDE - FL_H + lane;
  const int j0 = c * L;
  const bool out_lane = lane >= HL && lane < 64 - HL;
  const bool cok = gi >= 0 && gi < ni;
  const int ci = min(max(gi, 0), ni - 1);
  // raw rows j0 - 4 .. j0 + L + 3 (outside the block: 0), all requested here
  // from clamped addresses so the sequence has no branch
  T xr[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int gj = j0 - FL_H + r;
    if constexpr (INNER) {
      xr[r] = src[(size_t)gj * ni + gi];
    } else {
      const T v = src[(size_t)min(max(gj, 0), nj - 1) * ni + ci];
      xr[r] = (gj >= 0 && gj < nj && cok) ? v : (T)0;
    }
  }
  const int cx = cok ? default_code(gi, ni, NB) : CODE_ZERO;
  T xi[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    // w: columns gi - 4 .. gi + 4 of row r from lanes l - 4 .. l + 4
    T w[2 * FL_H + 1];
    T tl = xr[r], tr = xr[r];
    w[FL_H] = xr[r];
#pragma unroll
    for (int k = 1; k <= FL_H; k++) {
      tl = wave_shift1<true>(tl);
      tr = wave_shift1<false>(tr);
      w[FL_H - k] = tl;
      w[FL_H + k] = tr;
    }
    const int gj = j0 - FL_H + r;
    if constexpr (INNER) {
      xi[r] = filter_row<T, ORDER>(&w[FL_H], 1, CODE_INTERIOR);
    } else {
      const bool ok = gj >= 0 && gj < nj && cx != CODE_ZERO;
      xi[r] = ok ? filter_row<T, ORDER>(&w[FL_H], 1, cx) : (T)0;
    }
    if (r < 2 * FL_H) continue;
    // eta sweep, blend and store of output row o
    const int o = r - 2 * FL_H, go = j0 + o;
    if (!INNER && go >= nj) continue;
    const int cy = INNER ? CODE_INTERIOR : default_code(go, nj, NB);
    const T f = filter_row<T, ORDER>(&xi[o + FL_H], 1, cy);
    const T out = filter_blend(a.rem, xr[o + FL_H], a.blend, f);
    if (out_lane && (INNER || cok)) dst[(size_t)go * ni + (INNER ? gi : ci)] = out;
  }
}

// Stacked form of the column march: the NW waves of a workgroup filter one
// variable on one strip (64 lanes, 56 output columns) over NW R consecutive
// rows, R rows each.  Each wave computes the xi rows of its own R rows and
// 8 / NW of the 8 halo rows of the stack (4 above, 4 below), passes them to
// its neighbours through LDS, and takes the 4 xi rows on each side of its
// range from there; so a stack filters NW R + 8 xi rows instead of
// NW (R + 8).
template <typename T, int NW, int R>
struct StackLds {
  T xi[NW * R + 2 * FL_H][64];
};

// HOLES: per-point filter codes (xi at code[p], eta at code[n + p]) and the
// hole reset, as filter_tile's; INNER (never with HOLES): every load inside
// the block and every point of the stack on the interior rows.
template <typename T, int ORDER, int NW, int R, bool HOLES, bool INNER, typename AT>
__device__ __forceinline__ void stack_wave(const AT &a, int q, int s, int j0, int w, StackLds<T, NW, R> &L) {
  static_assert(!(HOLES && INNER), "an interior stack has no special points");
  constexpr int NB = ORDER / 2, HL = March::HL, NH = 2 * FL_H / NW;  // halo xi rows per wave
  static_assert(NW * NH == 2 * FL_H && R >= FL_H, "stack shape");
  const int lane = threadIdx.x;
  const int ni = a.ni, nj = a.nj;
  const size_t vs = a.vs ? a.vs : (size_t)ni * nj;
  const T *src = a.src + a.vars[q] * vs;
  T *dst = a.dst + a.vars[q] * vs;
  const int gi = s * March::STRIDE - FL_H + lane;
  const bool out_lane = lane >= HL && lane < 64 - HL;
  const bool cok = gi >= 0 && gi < ni;
  const int ci = min(max(gi, 0), ni - 1);
  const int cx = cok ? default_code(gi, ni, NB) : CODE_ZERO;
  // stack rows (LDS index): t = block row - (j0 - 4); own rows 4 + w R + k,
  // halo rows h = NH w .. NH w + NH - 1 of {0, .., 3, NW R + 4, .., NW R + 7}
  int tr[R + NH];
#pragma unroll
  for (int k = 0; k < R; k++) tr[k] = FL_H + w * R + k;
#pragma unroll
  for (int k = 0; k < NH; k++) {
    const int h = NH * w + k;
    tr[R + k] = h < FL_H ? h : NW * R + h;
  }
  T raw[R + NH];
  int8_t cxr[HOLES ? R + NH : 1], cyr[HOLES ? R : 1], hbr[HOLES ? R : 1];  // codes and iblank (HOLES)
#pragma unroll
  for (int k = 0; k < R + NH; k++) {
    const int gj = j0 - FL_H + tr[k];
    if constexpr (INNER) {
      raw[k] = src[(size_t)gj * ni + gi];
    } else {
      const size_t pc = (size_t)min(max(gj, 0), nj - 1) * ni + ci;
      raw[k] = src[pc];
      if constexpr (HOLES) {
        cxr[k] = a.code[pc];
        if (k < R) {
          cyr[k] = a.code[(size_t)ni * nj + pc];
          hbr[k] = a.iblank[pc];
        }
      }
    }
  }
  // xi rows (values outside the block are 0, selected here at the use)
  T xo[R + NH];
#pragma unroll
  for (int k = 0; k < R + NH; k++) {
    const int gj = j0 - FL_H + tr[k];
    const bool rok = INNER || (gj >= 0 && gj < nj);
    const T u = (rok && (INNER || cok)) ? raw[k] : (T)0;
    T wv[2 * FL_H + 1];
    T tl = u, trr = u;
    wv[FL_H] = u;
#pragma unroll
    for (int m = 1; m <= FL_H; m++) {
      tl = wave_shift1<true>(tl);
      trr = wave_shift1<false>(trr);
      wv[FL_H - m] = tl;
      wv[FL_H + m] = trr;
    }
    if constexpr (INNER) {
      xo[k] = filter_row<T, ORDER>(&wv[FL_H], 1, CODE_INTERIOR);
    } else {
      const int code = HOLES ? (cok ? (int)cxr[k] : CODE_ZERO) : cx;
      xo[k] = (rok && code != CODE_ZERO) ? filter_row<T, ORDER>(&wv[FL_H], 1, code) : (T)0;
    }
  }
  // neighbours need the first and last 4 own rows and the halo rows
#pragma unroll
  for (int k = 0; k < R + NH; k++)
    if (k >= R || k < FL_H || k >= R - FL_H) L.xi[tr[k]][lane] = xo[k];
  __syncthreads();
  T above[FL_H], below[FL_H];
#pragma unroll
  for (int m = 0; m < FL_H; m++) {
    above[m] = L.xi[w * R + m][lane];
    below[m] = L.xi[FL_H + (w + 1) * R + m][lane];
  }
#pragma unroll
  for (int k = 0; k < R; k++) {
    const int go = j0 + w * R + k;
    if (!INNER && go >= nj) break;
    T col[2 * FL_H + 1];
#pragma unroll
    for (int m = 0; m <= 2 * FL_H; m++) {
      const int r = k - FL_H + m;  // own-row index of the window row
      col[m] = r < 0 ? above[r + FL_H] : (r >= R ? below[r - R] : xo[r]);
    }
    const int cy = INNER ? CODE_INTERIOR : (HOLES ? (int)cyr[k] : default_code(go, nj, NB));
    T out = filter_blend(a.rem, raw[k], a.blend, filter_row<T, ORDER>(&col[FL_H], 1, cy));
    if (HOLES && hbr[k] == BLANKED) out = 0;
    if (out_lane && (INNER || cok)) dst[(size_t)go * ni + (INNER ? gi : ci)] = out;
  }
}

// Stack t of a launch: a.ntx strips, a.nty stacks of NW R rows per strip,
// a.ntiles = nvar * ntx * nty stacks.
template <typename T, int ORDER, int NW, int R, bool HOLES = false, typename AT>
__device__ __forceinline__ void stack_item(const AT &a, int t, StackLds<T, NW, R> &L) {
  const int c = t % a.nty, s = (t / a.nty) % a.ntx, q = t / (a.nty * a.ntx);
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int j0 = c * NW * R, i_s = s * March::STRIDE - FL_H;
  const bool inside = i_s >= 0 && i_s + 64 <= a.ni && j0 - FL_H >= 0 && j0 + NW * R + FL_H <= a.nj;
  if constexpr (HOLES) {
    // a stack whose output columns and xi rows hold no point off the
    // interior filter (no hole, no closure row) runs the interior instance
    if (inside && a.clean &&
        box_clean(a.clean, a.ni, a.nj, i_s + FL_H, i_s + 64 - FL_H, j0 - FL_H, j0 + NW * R + FL_H))
      stack_wave<T, ORDER, NW, R, false, true>(a, q, s, j0, w, L);
    else
      stack_wave<T, ORDER, NW, R, true, false>(a, q, s, j0, w, L);
  } else {
    if (inside)
      stack_wave<T, ORDER, NW, R, false, true>(a, q, s, j0, w, L);
    else
      stack_wave<T, ORDER, NW, R, false, false>(a, q, s, j0, w, L);
  }
}

// One stack per workgroup, a.chunk workgroups per XCD.
template <typename T, int ORDER, int NW, int R, bool HOLES>
__device__ __forceinline__ void filter_stack(const FilterArgs<T> &a, StackLds<T, NW, R> &L) {
  const int t = xcd_tile(blockIdx.x, a.chunk);
  if (t >= a.ntiles) return;
  stack_item<T, ORDER, NW, R, HOLES>(a, t, L);
}

// The wave's share from its index within the launch: four consecutive row
// chunks of one strip per workgroup, workgroups spread over the XCDs by
// xcd_tile so that each XCD walks whole strips (adjacent chunks share their
// halo rows in its L2).  a.ntx strips, a.nty chunks, a.ntiles = nvar * ntx *
// nty waves, a.chunk = workgroups per XCD.
template <typename T, int ORDER, int L>
__device__ __forceinline__ void filter_march(const FilterArgs<T> &a) {
  // the wave index is uniform: keep it (and every row index) in SGPRs
  const int wv = 4 * xcd_tile(blockIdx.x, a.chunk) + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  if (wv >= a.ntiles) return;
  const int c = wv % a.nty, s = (wv / a.nty) % a.ntx, q = wv / (a.nty * a.ntx);
  const int i_s = s * March::STRIDE - FL_H, j_s = c * L - FL_H;
  if (i_s >= 0 && i_s + 64 <= a.ni && j_s >= 0 && j_s + L + 2 * FL_H <= a.nj)
    march_wave<T, ORDER, L, true>(a, q, s, c);
  else
    march_wave<T, ORDER, L, false>(a, q, s, c);
}

}  // namespace

}  // namespace lee
