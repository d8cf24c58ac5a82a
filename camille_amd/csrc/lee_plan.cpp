// Host-side set-up for the LEE time advance: operator tables, per-point
// operator codes around holes, sponge strengths, the CFL candidate set and
// source supports.  All of it runs once in lee_create.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>

#include "lee_internal.h"

namespace lee {

// StandardCentral::initialize, ORDER_ACCURACY 4 (spatial_discretization.cpp:
// 28-109): interior {+-1: +-8/12, +-2: -+1/12}; closure rows i=0 {-3/2, 2,
// -1/2}, i=1 {-1/2, 0, 1/2}; mirrored with a sign flip on the right.
// OptimizedCentralExplicit::initialize, FDO11P (spatial_discretization.cpp:
// 228-344): Bogey-Bailly FDo11p interior; closure rows i=0, 1 Berland's FD06
// and FD15 (7 points), i=2 the standard 5-point row, i=3 FDo7p, i=4 FDo9p,
// all anchored at the boundary, 9 entries per row; right side mirrored with
// a sign flip.
static void make_fdo11p_table(OpTable &t) {
  t = OpTable{};
  t.half = 5;
  t.nb = 5;
  t.sb = 9;
  t.sign_right = -1.0;
  double *c = t.st + t.half;
  const double in[6] = {0.0, 0.872756993962, -0.286511173973, 0.090320001280, -0.020779405824, 0.002484594688};
  for (int s = 0; s <= 5; s++) c[s] = in[s];
  for (int s = 1; s <= 5; s++) c[-s] = -c[s];
  const double r0[7] = {-2.225833963270, 4.827779580575, -5.001388453836, 3.911103941646, -2.115267458633,
                        0.718882784412, -0.115276430894};
  const double r1[7] = {-0.212932721951, -1.060320390770, 2.078926116439, -1.287179452384, 0.685176395471,
                        -0.245320613994, 0.041650667189};
  const double r2[5] = {1.0 / 12.0, -8.0 / 12.0, 0.0, 8.0 / 12.0, -1.0 / 12.0};
  const double r3[7] = {-0.02625099337517627, 0.18833730683403838, -0.797921633542548, 0.0,
                        0.797921633542548,    -0.18833730683403838, 0.02625099337517627};
  const double r4[9] = {0.007650904064, -0.059463584768, 0.244678631765, -0.841570125482, 0.0,
                        0.841570125482, -0.244678631765, 0.059463584768, -0.007650904064};
  for (int s = 0; s < 7; s++) {
    t.bnd[0][s] = r0[s];
    t.bnd[1][s] = r1[s];
    t.bnd[3][s] = r3[s];
  }
  for (int s = 0; s < 5; s++) t.bnd[2][s] = r2[s];
  for (int s = 0; s < 9; s++) t.bnd[4][s] = r4[s];
}

WideOp make_wide_op(const OpTable &t) {
  WideOp w{};
  for (int s = 0; s < 11; s++) w.st[s] = t.st[s];
  for (int k = 0; k < 5; k++) {
    int len = 0;
    for (int s = 0; s < 9; s++) {
      w.bnd[k][s] = t.bnd[k][s];
      if (t.bnd[k][s] != 0.0) len = s + 1;
    }
    w.len[k] = len;
  }
  w.nb = t.nb;
  w.sign_right = t.sign_right;
  return w;
}

void make_fd_table(OpTable &t, int order) {
  if (order == 11) {
    make_fdo11p_table(t);
    return;
  }
  t = OpTable{};
  t.half = 2;
  t.nb = 2;
  t.sb = 3;
  t.sign_right = -1.0;
  double *c = t.st + t.half;
  c[0] = 0.0;
  c[1] = 8.0 / 12.0;
  c[2] = -1.0 / 12.0;
  c[-1] = -c[1];
  c[-2] = -c[2];
  t.bnd[0][0] = -3.0 / 2.0;
  t.bnd[0][1] = 4.0 / 2.0;
  t.bnd[0][2] = -1.0 / 2.0;
  t.bnd[1][0] = -1.0 / 2.0;
  t.bnd[1][1] = 0.0;
  t.bnd[1][2] = 1.0 / 2.0;
}

// StandardCentralFilter::initialize (spatial_discretization.cpp:113-224) at
// full strength (the standard filter always runs at strength 1,
// input.cpp:50-53).  Order 8: Bogey-Bailly SFs9p interior written as
// 1 - d0 at the centre and -d_k off it; closure rows 0 and 1 identity, row 2
// Lele's 4th-order and row 3 his 6th-order centred filters.  Order 4: Lele's
// 4th-order centred filter with identity rows 0 and 1.  Right side mirrored.
// OptimizedCentralFilterExplicit::initialize, SFO11P (spatial_discretization.
// cpp:348-492): Bogey-Bailly SFo11p interior at strength s (1 - s d0 at the
// centre, -s d_k off it); closure rows 0 and 1 identity, 2 and 3 Lele's 4th-
// and 6th-order filters, row 4 the 9-point optimized filter at strength s.
static void make_sfo11p_table(OpTable &t, double strength) {
  t = OpTable{};
  t.half = 5;
  t.nb = 5;
  t.sb = 9;
  t.sign_right = 1.0;
  double *c = t.st + t.half;
  const double d[6] = {0.215044884112, -0.187772883589, 0.123755948787, -0.059227575576, 0.018721609157,
                       -0.002999540835};
  c[0] = 1.0 - strength * d[0];
  for (int s = 1; s <= 5; s++) c[s] = d[s] * -strength;
  for (int s = 1; s <= 5; s++) c[-s] = c[s];
  t.bnd[0][0] = 1.0;
  t.bnd[1][1] = 1.0;
  const double lele4[5] = {(-1.0 / 8.0) / 2.0, (1.0 / 2.0) / 2.0, 5.0 / 8.0, (1.0 / 2.0) / 2.0, (-1.0 / 8.0) / 2.0};
  const double lele6[7] = {(1.0 / 32.0) / 2.0, (-3.0 / 16.0) / 2.0, (15.0 / 32.0) / 2.0, 11.0 / 16.0,
                           (15.0 / 32.0) / 2.0, (-3.0 / 16.0) / 2.0, (1.0 / 32.0) / 2.0};
  for (int s = 0; s < 5; s++) t.bnd[2][s] = lele4[s];
  for (int s = 0; s < 7; s++) t.bnd[3][s] = lele6[s];
  const double r4[9] = {0.008228661760, -0.045211119360, 0.120007591680, -0.204788880640, 0.243527493120,
                        -0.204788880640, 0.120007591680, -0.045211119360, 0.008228661760};
  for (int s = 0; s < 9; s++) t.bnd[4][s] = s == 4 ? 1.0 - strength * r4[s] : r4[s] * -strength;
}

void make_filter_table(OpTable &t, int order, double strength) {
  if (order == 11) {
    make_sfo11p_table(t, strength);
    return;
  }
  t = OpTable{};
  t.half = order / 2;
  t.nb = order / 2;
  t.sb = order - 1;
  t.sign_right = 1.0;
  double *c = t.st + t.half;
  if (order == 4) {
    c[0] = 5.0 / 8.0;
    c[1] = (1.0 / 2.0) / 2.0;
    c[2] = (-1.0 / 8.0) / 2.0;
    t.bnd[0][0] = 1.0;
    t.bnd[1][1] = 1.0;
  } else {
    const double d[5] = {35.0 / 128.0, -7.0 / 32.0, 7.0 / 64.0, -1.0 / 32.0, 1.0 / 256.0};
    const double strength = 1.0;
    c[0] = 1.0 - strength * d[0];
    for (int s = 1; s <= 4; s++) c[s] = d[s] * -strength;
    t.bnd[0][0] = 1.0;
    t.bnd[1][1] = 1.0;
    const double lele4[5] = {(-1.0 / 8.0) / 2.0, (1.0 / 2.0) / 2.0, 5.0 / 8.0, (1.0 / 2.0) / 2.0,
                             (-1.0 / 8.0) / 2.0};
    const double lele6[7] = {(1.0 / 32.0) / 2.0, (-3.0 / 16.0) / 2.0, (15.0 / 32.0) / 2.0, 11.0 / 16.0,
                             (15.0 / 32.0) / 2.0, (-3.0 / 16.0) / 2.0, (1.0 / 32.0) / 2.0};
    for (int s = 0; s < 5; s++) t.bnd[2][s] = lele4[s];
    for (int s = 0; s < 7; s++) t.bnd[3][s] = lele6[s];
  }
  for (int s = 1; s <= t.half; s++) c[-s] = c[s];
}

// Operator codes for one direction.  The reference first writes the
// interior rows and the block-edge closure rows (apply_spatial_discretization,
// spatial:1078-1290; extents geometry.cpp:931-1001), then overwrites cells
// around holes while scanning each grid line (iblank_correction,
// spatial:1356-1652): hole after hole -> 0; a fluid cell after a hole ->
// left-closure rows from there; a hole after a cell with iblank != 0 (the
// reference tests `last_iblank /= BLANKED`, which is non-zero only for an
// INTERPOLATED cell) -> right-closure rows ending at that cell.  Each write
// recomputes the cell from scratch, so the last one decides the code.
bool plan_codes(const Block &b, const OpTable &op, int dir, int8_t *codes, std::string &err) {
  const int nd[3] = {b.ni, b.nj, b.nk};
  const int nb = op.nb;
  const int len = nd[dir];
  const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
  auto lid = [&](int c, int c1, int c2) {
    int q[3];
    q[dir] = c;
    q[o1] = c1;
    q[o2] = c2;
    return (q[2] * b.nj + q[1]) * b.ni + q[0];
  };
  for (int c2 = 0; c2 < nd[o2]; c2++)
    for (int c1 = 0; c1 < nd[o1]; c1++) {
      for (int c = 0; c < len; c++) {
        int8_t code = CODE_INTERIOR;
        if (len > 1) {
          if (c < nb) code = (int8_t)(CODE_LEFT + c);
          else if (c > len - 1 - nb) code = (int8_t)(CODE_LEFT + nb + (len - 1 - c));
        }
        codes[lid(c, c1, c2)] = code;
      }
      if (!b.has_holes || len == 1) continue;
      auto ibk = [&](int c) { return (int)b.h_iblank[lid(c, c1, c2)]; };
      int last = ibk(0);
      for (int c = 0; c < len; c++) {
        int my = ibk(c);
        if (my == BLANKED) {
          if (last == BLANKED) {
            codes[lid(c, c1, c2)] = CODE_ZERO;
          } else if (last != 0 && c > 0) {
            int sum = 0;
            for (int k = 0; k < nb; k++) {
              int q = c - 1 - k;
              if (q < 0) { err = "hole closure would leave the block"; return false; }
              sum += std::max(ibk(q), 0);
            }
            if (sum != NOTBLANKED) { err = "too few fluid points between holes"; return false; }
            for (int k = 0; k < nb; k++) {
              int q = c - 1 - k;
              if (q >= 0 && q <= len - 1) codes[lid(q, c1, c2)] = (int8_t)(CODE_LEFT + nb + k);
            }
          }
        } else if (last == BLANKED) {
          int sum = 0;
          for (int k = 0; k < nb; k++) {
            int q = c + k;
            if (q >= len) { err = "hole closure would leave the block"; return false; }
            sum += std::max(ibk(q), 0);
          }
          if (sum != NOTBLANKED) { err = "too few fluid points between holes"; return false; }
          for (int k = 0; k < nb; k++) {
            int q = c + k;
            if (q >= 0 && q <= len - 1) codes[lid(q, c1, c2)] = (int8_t)(CODE_LEFT + k);
          }
        }
        last = my;
      }
    }
  return true;
}

// additionalInit_bufferZones (geometry.cpp:1028-1142): for every zone
// point, sigma = C (d_in / (d_in + d_out))^order, with d the minimum
// Euclidean distance to the zone's inflow face (away from the block
// boundary) and its outflow face (on it).  Exhaustive over the face points,
// split across host threads.
static void sponge_worker(const Block *b, const double *xyz, SpongeZone *z, int dir, int end, int order,
                          double constant, int k0, int k1) {
  const int n = b->n;
  const int fin = end == 0 ? z->ie[dir] : z->is[dir];
  const int fout = end == 0 ? z->is[dir] : z->ie[dir];
  const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
  // gather the two faces
  std::vector<double> face[2][3];
  for (int side = 0; side < 2; side++) {
    const int f = side == 0 ? fin : fout;
    for (int c2 = z->is[o2]; c2 <= z->ie[o2]; c2++)
      for (int c1 = z->is[o1]; c1 <= z->ie[o1]; c1++) {
        int q[3];
        q[dir] = f;
        q[o1] = c1;
        q[o2] = c2;
        int l = (q[2] * b->nj + q[1]) * b->ni + q[0];
        for (int d = 0; d < 3; d++) face[side][d].push_back(xyz[(size_t)d * n + l]);
      }
  }
  const int nf = (int)face[0][0].size();
  for (int lz = k0; lz < k1; lz++) {
    const int i = lz % z->nz[0], j = (lz / z->nz[0]) % z->nz[1], k = lz / (z->nz[0] * z->nz[1]);
    const int l0 = ((k + z->is[2]) * b->nj + (j + z->is[1])) * b->ni + (i + z->is[0]);
    const double px = xyz[l0], py = xyz[(size_t)n + l0], pz = xyz[(size_t)2 * n + l0];
    double dist2[2];
    for (int side = 0; side < 2; side++) {
      double dmin = 1.0e9;  // abs(DUMMY_DOUBLE) through int abs
      const double *fx = face[side][0].data(), *fy = face[side][1].data(), *fz = face[side][2].data();
      for (int q = 0; q < nf; q++) {
        double dd = 0.0;
        dd += std::pow(px - fx[q], 2);
        dd += std::pow(py - fy[q], 2);
        dd += std::pow(pz - fz[q], 2);
        dmin = std::min(dmin, dd);
      }
      dist2[side] = std::sqrt(dmin);
    }
    const double rel = dist2[0] / (dist2[0] + dist2[1]);
    z->host_sigma[lz] = constant * std::pow(rel, (double)order);
  }
}

void compute_sponge_zone(const Block &b, const double *xyz, SpongeZone &z, int dir, int end, int order,
                         double constant) {
  for (int d = 0; d < 3; d++) z.nz[d] = z.ie[d] - z.is[d] + 1;
  const int cnt = z.nz[0] * z.nz[1] * z.nz[2];
  z.host_sigma.assign(cnt, 0.0);
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (cnt < 4096) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++) {
    int k0 = (int)((long long)cnt * t / nt), k1 = (int)((long long)cnt * (t + 1) / nt);
    pool.emplace_back(sponge_worker, &b, xyz, &z, dir, end, order, constant, k0, k1);
  }
  for (auto &th : pool) th.join();
}

// temporal::compute_CFL (temporal_discretization.cpp:397-450).  Per point
// cfl = X * (dt * J) with
//   X = sum_{i,j} |trunc(U_i m_ij)|  (C int abs, :425; U from the mean flow)
//       + sqrt((gamma-1) T_bar) sqrt(sum m_ij^2),
// maximised over NOTBLANKED points.  X and J never change, so only pairs
// whose exact product X*J is within 1e-12 of the largest can attain the
// maximum of the rounded products for any dt (each side carries at most two
// roundings, 4.4e-16).  Keeping those distinct pairs reproduces the
// reference's maximum bit for bit at a cost independent of the grid.
void plan_cfl(Block &b, const double *mean, double gamma) {
  const int n = b.n;
  const double gm1 = gamma - 1.0;
  const double gi = 1.0 / gamma, gg = gamma / (gamma - 1.0);
  std::vector<double> X(n), prod(n);
  double pmax = -1e300;
  for (int p = 0; p < n; p++) {
    if (!b.h_iblank.empty() && b.h_iblank[p] != NOTBLANKED) {
      prod[p] = -1e300;
      continue;
    }
    // contravariant mean velocity and the metric sums over num_dim directions
    const int D = b.nk > 1 ? 3 : 2;
    double m[3][3];
    for (int l = 0; l < D; l++)
      for (int i = 0; i < D; i++) m[l][i] = b.h_met[(size_t)(D * l + i) * n + p];
    double uvec[3] = {mean[(size_t)IUX * n + p], mean[(size_t)IUY * n + p], mean[(size_t)IUZ * n + p]};
    double vk[3];
    for (int l = 0; l < D; l++) {
      vk[l] = 0.0;
      for (int i = 0; i < D; i++) vk[l] += m[l][i] * uvec[i];
    }
    double cfl = 0.0;
    for (int l = 0; l < D; l++)
      for (int i = 0; i < D; i++) cfl += std::abs((int)(vk[l] * m[l][i]));
    double fac = 0.0;
    for (int l = 0; l < D; l++)
      for (int i = 0; i < D; i++) fac += m[l][i] * m[l][i];
    const double sbar = mean[(size_t)IS * n + p], pbar = mean[(size_t)IP * n + p];
    const double rhobar = std::pow(gamma * pbar, gi) * std::exp(-sbar);
    const double Tbar = pbar * gg * (1.0 / rhobar);
    const double sos = std::sqrt(gm1 * Tbar);
    cfl += sos * std::sqrt(fac);
    X[p] = cfl;
    prod[p] = cfl * b.h_jac[p];
    pmax = std::max(pmax, prod[p]);
  }
  b.cfl_cand.clear();
  if (pmax <= -1e299) return;
  const double thr = pmax - std::fabs(pmax) * 1e-12;
  std::vector<std::pair<double, double>> c;
  for (int p = 0; p < n; p++)
    if (prod[p] >= thr) c.emplace_back(X[p], b.h_jac[p]);
  std::sort(c.begin(), c.end());
  c.erase(std::unique(c.begin(), c.end()), c.end());
  b.cfl_cand = c;
}

double cfl_max_block(const Block &b, double dt, double m) {
  for (const auto &c : b.cfl_cand) {
    double cfl = c.first;
    cfl *= dt * c.second;
    m = std::max(m, cfl);
  }
  return m;
}

double cfl_max(const lee_handle *h) {
  double m = -1e99;  // -DUMMY_LARGE
  for (int ib = 0; ib < h->nblk; ib++) m = cfl_max_block(h->b[ib], h->dt, m);
  return m;
}

// Physical sources (source.cpp:22-107) as spatial factor g(x) times time
// factors: term = (s1 * g) * s2, in the reference's multiplication order.
static double source_shape(const lee_config &c, double x, double y, double z) {
  switch (c.source_kind) {
    case LEE_SOURCE_TWOCYLINDER: {
      const double ef = -std::log(2.0) / (0.2 * 0.2);
      const double xs = x - 0.0, ys = y - 0.0;
      return std::exp(ef * (xs * xs + ys * ys));
    }
    case LEE_SOURCE_JET2D: {
      double Bx = 0.04 * std::log(2.0), By = 0.32 * std::log(2.0);
      const double bref = 1.3;
      Bx *= std::pow(bref, 2);
      By *= std::pow(bref, 2);
      return std::exp(-(Bx * x * x + By * y * y));
    }
    case LEE_SOURCE_MONOPOLE: {
      const double *q = c.source_params;
      const double k = -std::log(2.0) / (q[5] * q[5]);
      const double dx = x - q[2], dy = y - q[3], dz = z - q[4];
      return std::exp(k * (dx * dx + dy * dy + dz * dz));
    }
    default:
      return 0.0;
  }
}

void plan_source(const lee_config &c, Block &b, const double *xyz) {
  b.src.host_g.clear();
  b.src.i0 = 0;
  b.src.i1 = -1;
  b.src.j0 = 0;
  b.src.j1 = -1;
  b.src.k0 = 0;
  b.src.k1 = 0;
  if (c.source_kind == LEE_SOURCE_NONE) return;
  const int n = b.n;
  if (c.source_kind == LEE_SOURCE_FIELD) {  // the caller's field over the whole block, zero until set
    b.src.i0 = 0;
    b.src.i1 = b.ni - 1;
    b.src.j0 = 0;
    b.src.j1 = b.nj - 1;
    b.src.k1 = b.nk - 1;
    b.src.host_g.assign(n, 0.0);
    return;
  }
  std::vector<double> g(n);
  int i0 = b.ni, i1 = -1, j0 = b.nj, j1 = -1, k0 = b.nk, k1 = -1;
  for (int k = 0; k < b.nk; k++)
    for (int j = 0; j < b.nj; j++)
      for (int i = 0; i < b.ni; i++) {
        const int p = (k * b.nj + j) * b.ni + i;
        g[p] = source_shape(c, xyz[p], xyz[(size_t)n + p], xyz[(size_t)2 * n + p]);
        if (g[p] != 0.0) {
          i0 = std::min(i0, i);
          i1 = std::max(i1, i);
          j0 = std::min(j0, j);
          j1 = std::max(j1, j);
          k0 = std::min(k0, k);
          k1 = std::max(k1, k);
        }
      }
  if (i1 < i0) return;
  b.src.i0 = i0;
  b.src.i1 = i1;
  b.src.j0 = j0;
  b.src.j1 = j1;
  b.src.k0 = k0;
  b.src.k1 = k1;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) b.src.host_g.push_back(g[(k * b.nj + j) * b.ni + i]);
}

// bc_dirichlet_harmonicwave (boundary_condition.cpp:232-344) for LEE with
// WAVE_ENTROPY, the one wave type the reference's deck check lets through
// (input.cpp:636-648): s' = A (1 + sin(k c0 t)) with k = twopi / wavelength
// (twopi = 2 acos(-1), math.cpp:7-8, c0 = 1), times, for the round-jet
// shape, exp(-ln2 / halfwidth^2 (t1^2 + t2^2)) over the two coordinates
// transverse to the propagation direction (param.h:133).  Evaluated in the
// reference's operation order, on the host, so the device only multiplies.
double harmonic_entropy(const lee_config &c, double t) {
  const double twopi = 2.0 * std::acos(-1.0);
  const double wavenumber = twopi / c.harmonic_wavelength;
  const double c_0 = 1.0;
  return c.harmonic_amplitude * (1.0 + std::sin(wavenumber * c_0 * t));
}

double harmonic_gaussian(const lee_config &c, double x, double y, double z) {
  static const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
  const double xyz[3] = {x, y, z};
  const int d = c.harmonic_direction;
  const double t1 = xyz[other[d][0]], t2 = xyz[other[d][1]];
  const double fac = -std::log(2.0) / std::pow(c.harmonic_halfwidth, 2);
  return std::exp(fac * (std::pow(t1, 2) + std::pow(t2, 2)));
}

void source_factors(const lee_config &c, double t, double &s1, double &s2) {
  const double pi = std::acos(-1.0), twopi = 2.0 * pi;
  s1 = 1.0;
  s2 = 0.0;
  switch (c.source_kind) {
    case LEE_SOURCE_TWOCYLINDER: {
      const double mf = c.gamma - 1.0;
      const double sin_wt = std::sin(4.0 * twopi * t);
      double ramp = 1.0;
      const double ramp_time = 16.0 / 4.0;
      if (t < ramp_time) ramp = std::pow(std::sin(pi / 2.0 * t / ramp_time), 2);
      s1 = ramp * mf;
      s2 = sin_wt;
      break;
    }
    case LEE_SOURCE_JET2D: {
      double amplitude = 0.001, frequency = 76.0;
      const double bref = 1.3, R = 287.0, rhoinf = 1.184, Tinf = 300.0;
      const double cinf = std::sqrt(c.gamma * R * Tinf);
      amplitude /= rhoinf * std::pow(cinf, 2) / (bref / cinf);
      frequency *= (1.0 / twopi) * (bref / cinf);
      const double w = frequency * twopi;
      s2 = amplitude * std::cos(w * t);
      break;
    }
    case LEE_SOURCE_MONOPOLE:
    case LEE_SOURCE_FIELD:
      s2 = c.source_params[0] * std::sin(c.source_params[1] * t);
      break;
    default:
      break;
  }
}

}  // namespace lee
