// BP-gauged two-site gate application (simple update) on the device.
//
// Semantics: ITensors.apply(o, psi; envs, maxdim, cutoff, normalize) with a
// product environment (src/apply.jl:97-146) -> simple_update_bp (src/apply.jl:33-95):
//   1. every environment message M (chi x chi, Hermitian) -> sqrt(M) and M^{-1/2}
//      by a Hermitian eigendecomposition with the eigenvalue truncation
//      cutoff = 10 eps (map_eigvals, src/apply.jl:9-25);
//   2. psi_u~ = psi_u with sqrt(M) on every external bond, reshaped to
//      X_u (external bonds) x (s_u, b_uv), QR X_u = Q_u R_u (src/apply.jl:74-75);
//   3. Theta = gate * (R_u R_v) over b_uv (src/apply.jl:78), SVD, truncation to
//      maxdim / cutoff with truncerr = discarded weight (src/apply.jl:81-89);
//   4. psi_u' = (Q_u with M^{-1/2} on the external bonds) U sqrt(S), likewise
//      psi_v' with sqrt(S) V^T (src/apply.jl:90-93); optional normalisation
//      (src/apply.jl:134-137).
// One workgroup per bond; the bonds of one launch are disjoint.  Small dense
// linear algebra: two-sided Jacobi in registers for the environment
// eigenproblems (one wave each), Householder QR and one-sided (Hestenes) Jacobi
// SVD in LDS.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "itnbp_ops.hpp"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int AP_XY = 8192;   // max E * n (external size x d chi_b) per side
// tall-skinny QR (E >= 4 n, n <= 16): the stacked R factors (4n x n) and the
// heads and betas of their reflectors, per side, after the side's reflectors
constexpr int AP_QRX = 4 * 16 * 16 + 4 * 16;
constexpr int AP_SIDE = AP_XY + AP_QRX;  // global scratch per side of a gate
constexpr int AP_TH = 32;     // max rows/cols of Theta (r d)
constexpr int AP_ENV = 8;     // max external bond dimension
constexpr int AP_MAXEXT = 3;  // max external bonds per side (degree 4)
constexpr int AP_SITE_D = 16; // max physical dimension of a one-site gate
// the global-memory variant (k_apply_gates<true>) for bonds past the LDS limits:
// every per-bond matrix in a global workspace, one workgroup per bond
constexpr int AP_ENV_BIG = 32;    // max external bond dimension
constexpr int AP_N_BIG = 64;      // max d chi_b
constexpr int AP_TH_BIG = 128;    // max rows/cols of Theta
constexpr int AP_KEEP_BIG = 64;   // max kept rank (the growth cap of itnbp_apply_gates)
constexpr int AP_F_BIG = AP_TH_BIG * AP_KEEP_BIG;  // R factor, then the left / right factor
// LDS of the small vectors (rot, four reflector-head vectors, the singular
// values, the block reduction), rounded up to 256 bytes
constexpr int AP_BIG_SMALL = (4 * 32 + 4 * AP_N_BIG + AP_TH_BIG + 16 + 31) & ~31;

struct ApplyItem {
  int u, v;      // vertices
  int ju, jv;    // bond axis at u, v
  long long gate;  // offset of the (d_u d_v)^2 gate
  int slot;
  int chib_in;     // bond dimension before the call (the layout may be grown past it)
};

struct ApplyParams {
  int maxdim;         // <= 0: none (still capped by the bond dimension)
  double cutoff;      // < 0: none
  int normalize;
  double eig_cutoff;  // map_eigvals cutoff (10 eps)
  int max_keep;       // stride of the singular value output
  // per-phase timestamps (wall clock; 16 slots per gate, slot 8 the SVD's
  // sweep count), or null; ITNBP_APPLY_PHASES=1
  unsigned long long* ts = nullptr;
  // global-memory variant: per-gate workspace stride and the sizes of its
  // side matrices (QR of each side: E n; work: E max(n, d keep)), in doubles
  long long big_stride = 0, big_q = 0, big_x = 0;
};

// thread 0 of the workgroup records the end of phase ph (0: start)
__device__ inline void phase_mark(const ApplyParams& prm, int ph) {
  if (prm.ts && threadIdx.x == 0) prm.ts[blockIdx.x * 16 + ph] = wall_clock64();
}

__device__ inline double wave_sum_d(double v) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

__device__ double block_sum_d(double v, double* red) {
  v = wave_sum_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  __syncthreads();
  return s;
}

// Round-robin pairing: step st of a cyclic sweep over n (even) indices.
__device__ inline void rr_pair(int n, int st, int i, int& p, int& q) {
  // positions 0..n-1; position 0 fixed, others rotate
  // (pos - 1 + st) < 2 (n - 1): one conditional subtraction instead of a modulo
  auto at = [&](int pos) {
    if (pos == 0) return 0;
    const int x = pos - 1 + st;
    return 1 + (x >= n - 1 ? x - (n - 1) : x);
  };
  p = at(i);
  q = at(n - 1 - i);
  if (p > q) {
    int t = p;
    p = q;
    q = t;
  }
}

// sqrt(M) and M^{-1/2} of a symmetrised chi x chi message (column-major in
// global), chi <= 8, by one wave: lane i ne + j (ne = chi rounded up to even)
// holds A[i][j] and V[i][j] in registers.  Cyclic two-sided Jacobi with
// round-robin pairs (rr_pair's schedule), every lane deriving the rotations of
// its row pair and column pair from the pre-step matrix through shuffles, so a
// step needs no LDS and no barrier; stops when the off-diagonal mass is below
// 1e-30 of the diagonal's or stops decreasing (at most 40 sweeps).  Then the eigenvalue truncation
// of map_eigvals (cutoff = 10 eps, src/apply.jl:9-25), recomputed identically
// by every lane.  S and Si row-major, ld chi.
__device__ void env_roots_wave(const double* M, int chi, double cutoff, double* S, double* Si) {
  const int l = threadIdx.x & 63;
  const int ne = chi + (chi & 1);
  const bool live = l < ne * ne;
  const int i = live ? l / ne : 0, j = live ? l % ne : 0;
  const bool in = live && i < chi && j < chi;
  double a = in ? 0.5 * (M[i + chi * j] + M[j + chi * i]) : 0.0;
  double v = (live && i == j) ? 1.0 : 0.0;
  if (chi > 1) {
    auto rot = [&](int p, int q, double& c, double& sn) {
      const double app = __shfl(a, p * ne + p), aqq = __shfl(a, q * ne + q), apq = __shfl(a, p * ne + q);
      c = 1.0;
      sn = 0.0;
      if (q < chi && apq != 0.0) {
        const double zeta = (aqq - app) / (2 * apq);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        c = 1 / sqrt(1 + t * t);
        sn = c * t;
      }
    };
    double prev = INFINITY;
    for (int sweep = 0; sweep < 40; ++sweep) {
      double off = (in && i != j) ? a * a : 0.0, diag = (in && i == j) ? a * a : 0.0;
      for (int m = 32; m > 0; m >>= 1) {
        off += __shfl_xor(off, m);
        diag += __shfl_xor(diag, m);
      }
      // converged, or the off-diagonal mass no longer decreases (rounding level)
      if (off <= 1e-30 * diag || off == 0.0 || off >= prev) break;
      prev = off;
      for (int st = 0; st < ne - 1; ++st) {
        // partner of index x in step st of the round-robin schedule (rr_pair)
        // (x - 1 - st) lies in (-(ne - 1), ne - 1) and (pp - 1 + st) in [0, 2 (ne - 1)):
        // conditional corrections instead of modulos
        auto partner = [&](int x) {
          int t = x - 1 - st;
          if (t < 0) t += ne - 1;
          const int pos = x == 0 ? 0 : 1 + t;
          const int pp = ne - 1 - pos;
          int u = pp - 1 + st;
          if (u >= ne - 1) u -= ne - 1;
          return pp == 0 ? 0 : 1 + u;
        };
        const int ri = partner(i), cj = partner(j);
        double cr, sr, cc, sc;
        rot(i < ri ? i : ri, i < ri ? ri : i, cr, sr);
        rot(j < cj ? j : cj, j < cj ? cj : j, cc, sc);
        const double ao = __shfl(a, ri * ne + j);  // rows: A <- J^T A
        const double ar = i < ri ? cr * a - sr * ao : sr * ao + cr * a;
        const double aro = __shfl(ar, i * ne + cj), vo = __shfl(v, i * ne + cj);  // columns: A J, V J
        a = j < cj ? cc * ar - sc * aro : sc * aro + cc * ar;
        v = j < cj ? cc * v - sc * vo : sc * vo + cc * v;
      }
    }
  }
  double w[AP_ENV];
#pragma unroll
  for (int m = 0; m < AP_ENV; ++m) w[m] = __shfl(a, (m < ne ? m : 0) * ne + (m < ne ? m : 0));
  // truncation: rank |lambda| descending (ties by index); drop from the small
  // end while the dropped weight / total <= cutoff (keep >= 1).  Fully
  // unrolled over AP_ENV, with no run-time index into w, so that w stays in
  // registers.
  int rank[AP_ENV];
  double tot = 0;
#pragma unroll
  for (int m = 0; m < AP_ENV; ++m) {
    int r = 0;
#pragma unroll
    for (int m2 = 0; m2 < AP_ENV; ++m2)
      if (m2 < chi && (fabs(w[m2]) > fabs(w[m]) || (fabs(w[m2]) == fabs(w[m]) && m2 < m))) ++r;
    rank[m] = r;
    if (m < chi) tot += fabs(w[m]);
  }
  // the value of rank r, and the dropped prefix from the small end
  auto by_rank = [&](int r) {
    double x = 0;
#pragma unroll
    for (int m = 0; m < AP_ENV; ++m)
      if (m < chi && rank[m] == r) x = fabs(w[m]);
    return x;
  };
  int keep = chi;
  double disc = 0;
  while (keep > 1 && tot > 0 && (disc + by_rank(keep - 1)) / tot <= cutoff) {
    disc += by_rank(keep - 1);
    --keep;
  }
  unsigned kmask = 0;
#pragma unroll
  for (int m = 0; m < AP_ENV; ++m)
    if (m < chi && rank[m] < keep) kmask |= 1u << m;
  // S[r][c] = sum_m V[r][m] V[c][m] sqrt(l_m) over the kept m (lane (r, c) = (i, j))
  double sv = 0, siv = 0;
  for (int m = 0; m < AP_ENV; ++m) {
    const int mm = m < ne ? m : 0;
    const double vr = __shfl(v, i * ne + mm), vc = __shfl(v, j * ne + mm);
    if (m < chi && ((kmask >> m) & 1u)) {
      const double sq = sqrt(w[m]);  // NaN for a kept negative eigenvalue
      sv += vr * vc * sq;
      siv += vr * vc / sq;
    }
  }
  if (in) {
    S[i * chi + j] = sv;
    Si[i * chi + j] = siv;
  }
}

__device__ inline void wave_sync_ap() {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
}

// sqrt(M) and M^{-1/2} of a symmetrised chi x chi message for chi <= AP_ENV_BIG
// (the global-memory variant), by one wave with the matrices in LDS (wk:
// ENVB_WORK doubles per wave): cyclic two-sided Jacobi with the round-robin
// pairs of rr_pair, the rotations of a step derived by the first ne / 2 lanes
// and then applied to the rows, the columns and V; the stopping rule of
// env_roots_wave (off-diagonal mass below 1e-30 of the diagonal's, or no
// longer decreasing, at most 40 sweeps) and its eigenvalue truncation
// (map_eigvals, src/apply.jl:9-25).  S and Si row-major, ld chi.
constexpr int ENVB_LD = AP_ENV_BIG + 1;
constexpr int ENVB_WORK = 2 * AP_ENV_BIG * ENVB_LD + 4 * AP_ENV_BIG;
__device__ void env_roots_big(const double* M, int chi, double cutoff, double* S, double* Si, double* wk) {
  const int l = threadIdx.x & 63;
  double* A = wk;
  double* V = A + AP_ENV_BIG * ENVB_LD;
  double* rc = V + AP_ENV_BIG * ENVB_LD;  // per pair of a step: c, s, p, q
  double* rs = rc + AP_ENV_BIG / 2;
  int* rp = reinterpret_cast<int*>(rs + AP_ENV_BIG / 2);
  int* rq = rp + AP_ENV_BIG / 2;
  int* rank = rq + AP_ENV_BIG / 2;        // rank of each eigenvalue by |lambda|
  int* byr = rank + AP_ENV_BIG;           // the eigenvalue of each rank
  for (int o = l; o < chi * chi; o += 64) {
    const int i = o / chi, j = o - i * chi;
    A[i * ENVB_LD + j] = 0.5 * (M[i + chi * j] + M[j + chi * i]);
    V[i * ENVB_LD + j] = i == j ? 1.0 : 0.0;
  }
  wave_sync_ap();
  const int ne = chi + (chi & 1), np = ne / 2;
  double prev = INFINITY;
  for (int sweep = 0; sweep < 40 && chi > 1; ++sweep) {
    double off = 0, dg = 0;
    for (int o = l; o < chi * chi; o += 64) {
      const int i = o / chi, j = o - i * chi;
      const double a = A[i * ENVB_LD + j];
      if (i != j)
        off += a * a;
      else
        dg += a * a;
    }
    off = wave_sum_d(off);
    dg = wave_sum_d(dg);
    if (off <= 1e-30 * dg || off == 0.0 || off >= prev) break;
    prev = off;
    for (int st = 0; st < ne - 1; ++st) {
      if (l < np) {
        int p, q;
        rr_pair(ne, st, l, p, q);
        double c = 1, sn = 0;
        if (q < chi) {
          const double apq = A[p * ENVB_LD + q];
          if (apq != 0.0) {
            const double zeta = (A[q * ENVB_LD + q] - A[p * ENVB_LD + p]) / (2 * apq);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
            c = 1 / sqrt(1 + t * t);
            sn = c * t;
          }
        }
        rc[l] = c;
        rs[l] = sn;
        rp[l] = p;
        rq[l] = q;
      }
      wave_sync_ap();
      for (int o = l; o < np * chi; o += 64) {  // rows: A <- J^T A
        const int i = o / chi, col = o - i * chi;
        const int p = rp[i], q = rq[i];
        if (q >= chi) continue;
        const double c = rc[i], sn = rs[i];
        const double ap = A[p * ENVB_LD + col], aq = A[q * ENVB_LD + col];
        A[p * ENVB_LD + col] = c * ap - sn * aq;
        A[q * ENVB_LD + col] = sn * ap + c * aq;
      }
      wave_sync_ap();
      for (int o = l; o < np * chi; o += 64) {  // columns: A <- A J, V <- V J
        const int i = o / chi, row = o - i * chi;
        const int p = rp[i], q = rq[i];
        if (q >= chi) continue;
        const double c = rc[i], sn = rs[i];
        const double ap = A[row * ENVB_LD + p], aq = A[row * ENVB_LD + q];
        A[row * ENVB_LD + p] = c * ap - sn * aq;
        A[row * ENVB_LD + q] = sn * ap + c * aq;
        const double vp = V[row * ENVB_LD + p], vq = V[row * ENVB_LD + q];
        V[row * ENVB_LD + p] = c * vp - sn * vq;
        V[row * ENVB_LD + q] = sn * vp + c * vq;
      }
      wave_sync_ap();
    }
  }
  // truncation: rank |lambda| descending (ties by index); drop from the small
  // end while the dropped weight / total <= cutoff (keep >= 1)
  if (l < chi) {
    const double wl = fabs(A[l * ENVB_LD + l]);
    int r = 0;
    for (int m = 0; m < chi; ++m) {
      const double wm = fabs(A[m * ENVB_LD + m]);
      if (wm > wl || (wm == wl && m < l)) ++r;
    }
    rank[l] = r;
    byr[r] = l;
  }
  wave_sync_ap();
  double tot = 0;
  for (int m = 0; m < chi; ++m) tot += fabs(A[m * ENVB_LD + m]);
  int keep = chi;
  double disc = 0;
  while (keep > 1 && tot > 0) {
    const int m = byr[keep - 1];
    const double x = fabs(A[m * ENVB_LD + m]);
    if ((disc + x) / tot > cutoff) break;
    disc += x;
    --keep;
  }
  // S[i][j] = sum_m V[i][m] V[j][m] sqrt(l_m) over the kept m, Si with 1 / sqrt
  for (int o = l; o < chi * chi; o += 64) {
    const int i = o / chi, j = o - i * chi;
    double sv = 0, siv = 0;
    for (int m = 0; m < chi; ++m) {
      if (rank[m] >= keep) continue;
      const double sq = sqrt(A[m * ENVB_LD + m]);  // NaN for a kept negative eigenvalue
      const double vv = V[i * ENVB_LD + m] * V[j * ENVB_LD + m];
      sv += vv * sq;
      siv += vv / sq;
    }
    S[i * chi + j] = sv;
    Si[i * chi + j] = siv;
  }
  wave_sync_ap();
}

// Exact x / d and x % d by a multiply-high: m = ceil(2^sh / d) with
// sh = 32 + ceil(log2 d), q = (x m) >> sh, exact for every x < 2^31 (x d <
// 2^sh, and x m fits 64 bits); the integer divisions of the index arithmetic
// otherwise cost ~30 instructions each, and the flat element loops of this
// kernel run one wave per SIMD with nothing to hide them behind.
struct FastDiv {  // an aggregate, so that a SideGeom can live in LDS
  unsigned d, sh;
  unsigned long long m;
  __device__ unsigned div(unsigned x) const { return (unsigned)(((unsigned long long)x * m) >> sh); }
  __device__ unsigned mod(unsigned x) const { return x - div(x) * d; }
};
__device__ inline FastDiv fast_div(unsigned d) {
  unsigned sh = 32;
  while ((1ull << (sh - 32)) < d) ++sh;
  return FastDiv{d, sh, ((1ull << sh) + d - 1) / d};
}

struct SideGeom {
  int d, k, j, chib, E, n, r;
  int dims[ITNBP_MAXR];
  int ext[AP_MAXEXT];  // external axes in order
  int next;
  long long psi_off;
  FastDiv fE, fd, fdim[ITNBP_MAXR];
  // external axis pa: stride of its index in e, its dimension, and their product
  FastDiv fst[AP_MAXEXT], fchi[AP_MAXEXT];
};

__device__ SideGeom side_geom(const VDesc& V, int j) {
  SideGeom g;
  g.d = V.dims[0];
  g.k = V.k;
  g.j = j;
  g.chib = V.dims[j];
  g.E = 1;
  g.next = 0;
  for (int a = 0; a <= V.k; ++a) g.dims[a] = V.dims[a];
  for (int a = 1; a <= V.k; ++a)
    if (a != j) {
      g.ext[g.next++] = a;
      g.E *= V.dims[a];
    }
  g.n = g.d * g.chib;
  g.r = g.E < g.n ? g.E : g.n;
  g.psi_off = V.psi_off;
  g.fE = fast_div(g.E);
  g.fd = fast_div(g.d);
  for (int a = 0; a < ITNBP_MAXR; ++a) g.fdim[a] = fast_div(a <= V.k ? V.dims[a] : 1);
  int st = 1;
  for (int pa = 0; pa < AP_MAXEXT; ++pa) {
    const int chi = pa < g.next ? V.dims[g.ext[pa]] : 1;
    g.fst[pa] = fast_div(st);
    g.fchi[pa] = fast_div(chi);
    st *= chi;
  }
  return g;
}

// the tall-skinny QR applies: four row blocks of at least n rows, n <= 16
__device__ inline bool tsqr_ok(const SideGeom& g) { return g.n <= 16 && g.E >= 4 * g.n; }

// natural offset of psi element (s, external multi-index e, bond b)
__device__ inline int nat_off(const SideGeom& g, int s, int e, int b) {
  int off = 0, stride = g.d, rem = e;
  for (int a = 1; a <= g.k; ++a) {
    int x;
    if (a == g.j) {
      x = b;
    } else {
      const int q = (int)g.fdim[a].div(rem);
      x = rem - q * g.dims[a];
      rem = q;
    }
    off += stride * x;
    stride *= g.dims[a];
  }
  return s + off;
}

// mode product on external axis index pa of X (E x cols, column-major rows e):
// Y[e with b'] = sum_b X[e with b] * Mt[b][b'] (Mt row-major, ld chi), chi <= CH
template <int CH>
__device__ void ext_mode_product(const SideGeom& g, int pa, const double* X, double* Y, int cols,
                                 const double* Mt, bool transpose) {
  const int stride = (int)g.fst[pa].d;
  const int chi = g.dims[g.ext[pa]];
  for (int idx = threadIdx.x; idx < g.E * cols; idx += blockDim.x) {
    const int c = (int)g.fE.div(idx), e = idx - c * g.E;
    const int q = (int)g.fst[pa].div(e), lo = e - q * stride;
    const int hi = (int)g.fchi[pa].div(q), b2 = q - hi * chi;
    double acc = 0;
#pragma unroll
    for (int b = 0; b < CH; ++b)
      if (b < chi) {
        const double m = transpose ? Mt[b2 * chi + b] : Mt[b * chi + b2];
        acc += X[(lo + stride * (b + chi * hi)) + (long long)g.E * c] * m;
      }
    Y[idx] = acc;
  }
  __syncthreads();
}

// The same mode product for the global-memory variant, one fibre per thread:
// a thread loads the chi entries of one (lo, hi, column) fibre once and writes
// all chi outputs, with Mt (transposed if asked) staged in LDS at ms (CH x CH;
// every lane of a wave reads the same entry).  The per-output form above
// reloads each fibre chi times, which from global memory costs more than the
// multiply-adds.
template <int CH>
__device__ void ext_mode_product_fibre(const SideGeom& g, int pa, const double* X, double* Y, int cols,
                                       const double* Mt, bool transpose, double* ms) {
  const int stride = (int)g.fst[pa].d;
  const int chi = g.dims[g.ext[pa]];
  for (int i = threadIdx.x; i < chi * chi; i += blockDim.x) {
    const int b = i / chi, b2 = i - b * chi;
    ms[b * CH + b2] = transpose ? Mt[b2 * chi + b] : Mt[b * chi + b2];
  }
  __syncthreads();
  const int per = g.E / chi;  // fibres per column
  const FastDiv fper = fast_div((unsigned)per);
  const int nfib = per * cols;
  for (int f = threadIdx.x; f < nfib; f += blockDim.x) {
    const int c = (int)fper.div(f), r = f - c * per;
    const int hi = (int)g.fst[pa].div(r), lo = r - hi * stride;
    const long long base = (long long)g.E * c + lo + (long long)stride * chi * hi;
    double x[CH];
#pragma unroll
    for (int b = 0; b < CH; ++b) x[b] = b < chi ? X[base + (long long)stride * b] : 0.0;
#pragma unroll 4
    for (int b2 = 0; b2 < chi; ++b2) {
      double acc = 0;
#pragma unroll
      for (int b = 0; b < CH; ++b) acc += x[b] * ms[b * CH + b2];
      Y[base + (long long)stride * b2] = acc;
    }
  }
  __syncthreads();
}

// Householder QR in place on X (E x n, column-major); R in the upper triangle,
// reflector tails below, heads in vh[], betas in beta[]
__device__ void householder_qr(double* X, int E, int n, double* vh, double* beta, double* red) {
  const int tid = threadIdx.x;
  const int r = E < n ? E : n;
  for (int j = 0; j < r; ++j) {
    double s = 0;
    for (int i = j + tid; i < E; i += blockDim.x) s += X[i + (long long)E * j] * X[i + (long long)E * j];
    s = block_sum_d(s, red);
    const double xjj = X[j + (long long)E * j];
    const double nrm = sqrt(s);
    const double alpha = xjj >= 0 ? -nrm : nrm;
    const double v0 = xjj - alpha;
    const double vv = s - xjj * xjj + v0 * v0;
    const double b = (nrm == 0.0 || vv == 0.0) ? 0.0 : 2.0 / vv;
    if (tid == 0) {
      vh[j] = v0;
      beta[j] = b;
    }
    // apply to columns c > j: w = v . X[:, c]; X[:, c] -= b w v
    const int w = tid >> 6, l = tid & 63;
    for (int c = j + 1 + w; c < n; c += (int)(blockDim.x >> 6)) {
      double dot = 0;
      for (int i = j + l; i < E; i += 64) {
        const double vi = (i == j) ? v0 : X[i + (long long)E * j];
        dot += vi * X[i + (long long)E * c];
      }
      dot = wave_sum_d(dot);
      for (int i = j + l; i < E; i += 64) {
        const double vi = (i == j) ? v0 : X[i + (long long)E * j];
        X[i + (long long)E * c] -= b * dot * vi;
      }
    }
    __syncthreads();
    if (tid == 0) X[j + (long long)E * j] = (b == 0.0) ? xjj : alpha;
    __syncthreads();
  }
}

// Y (E x cols) <- H_0 ... H_{r-1} Y using the reflectors stored in X
__device__ void apply_q(const double* X, int E, int r, const double* vh, const double* beta, double* Y,
                        int cols) {
  const int tid = threadIdx.x;
  const int w = tid >> 6, l = tid & 63;
  for (int j = r - 1; j >= 0; --j) {
    const double b = beta[j];
    if (b != 0.0) {
      for (int c = w; c < cols; c += (int)(blockDim.x >> 6)) {
        double dot = 0;
        for (int i = j + l; i < E; i += 64) {
          const double vi = (i == j) ? vh[j] : X[i + (long long)E * j];
          dot += vi * Y[i + (long long)E * c];
        }
        dot = wave_sum_d(dot);
        for (int i = j + l; i < E; i += 64) {
          const double vi = (i == j) ? vh[j] : X[i + (long long)E * j];
          Y[i + (long long)E * c] -= b * dot * vi;
        }
      }
    }
    __syncthreads();
  }
}

// Tall-skinny QR of a side (E x n, E >= 4 n, n <= 16), replacing the
// workgroup-wide Householder QR there: each wave factors its quarter of the
// rows (rows [w E / 4, (w + 1) E / 4)) with no workgroup barrier, wave 0 then
// factors the four stacked n x n R factors (4n x n), whose R is the side's.
// Q = diag(Q_0 .. Q_3) P Q_S (P places stacked row w n + i at block row i of
// block w), so X = Q [R; 0] as for one Householder QR, up to the signs of R's
// rows, which the rebuild's Q absorbs.
//
// Householder QR of rows x n (n <= 16) at X (column stride ld) by one wave:
// lane (rg, c) = (l / 16, l % 16) works on rows i = rg (mod 4) of column c,
// partial sums joined across rg by two shuffles; reflector j is
// [vh[j]; X[j+1 .., j]] with beta[j], R in the upper triangle.  Every element
// is read and written by its own lane, so only the diagonal element (read by
// every lane) needs the wave to be in step.
__device__ void wave_qr16(double* X, int ld, int rows, int n, double* vh, double* beta) {
  const int l = threadIdx.x & 63, rg = l >> 4, c = l & 15;
  for (int j = 0; j < n; ++j) {
    const int i0 = j + 1 + ((rg - ((j + 1) & 3)) & 3);  // first row > j with i = rg (mod 4)
    // four rows per iteration with independent partial sums, so that the LDS
    // loads of an iteration are in flight together (one wave per SIMD here:
    // nothing else hides their latency)
    double p = 0;
    if (c < n) {
      const double* xj = X + (long long)ld * j;
      const double* xc = X + (long long)ld * c;
      double p0 = 0, p1 = 0, p2 = 0, p3 = 0;
      int i = i0;
      for (; i + 12 < rows; i += 16) {
        p0 += xj[i] * xc[i];
        p1 += xj[i + 4] * xc[i + 4];
        p2 += xj[i + 8] * xc[i + 8];
        p3 += xj[i + 12] * xc[i + 12];
      }
      for (; i < rows; i += 4) p0 += xj[i] * xc[i];
      p = (p0 + p1) + (p2 + p3);
    }
    p += __shfl_xor(p, 16);
    p += __shfl_xor(p, 32);
    const double tail = __shfl(p, j);  // sum_{i > j} X[i][j]^2
    const double xjj = X[j + (long long)ld * j];
    const double s = xjj * xjj + tail;
    const double nrm = sqrt(s);
    const double alpha = xjj >= 0 ? -nrm : nrm;
    const double v0 = xjj - alpha;
    const double vv = tail + v0 * v0;
    const double b = (nrm == 0.0 || vv == 0.0) ? 0.0 : 2.0 / vv;
    if (c > j && c < n) {
      const double dot = v0 * X[j + (long long)ld * c] + p;
      const double f = b * dot;
      if (rg == 0) X[j + (long long)ld * c] -= f * v0;
      const double* xj = X + (long long)ld * j;
      double* xc = X + (long long)ld * c;
      int i = i0;
      for (; i + 12 < rows; i += 16) {
        const double a0 = xj[i], a1 = xj[i + 4], a2 = xj[i + 8], a3 = xj[i + 12];
        const double c0 = xc[i], c1 = xc[i + 4], c2 = xc[i + 8], c3 = xc[i + 12];
        xc[i] = c0 - f * a0;
        xc[i + 4] = c1 - f * a1;
        xc[i + 8] = c2 - f * a2;
        xc[i + 12] = c3 - f * a3;
      }
      for (; i < rows; i += 4) xc[i] -= f * xj[i];
    }
    if (l == 0) {
      vh[j] = v0;
      beta[j] = b;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (l == 0) X[j + (long long)ld * j] = (b == 0.0) ? xjj : alpha;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
  }
}

// Y[rows(i), c] <- H_0 .. H_{n-1} Y for c < cols (cols <= 16 per pass) by one
// wave: reflector j is [vh[j]; V[j+1 .., j]] (column stride ldv) over the
// logical rows i >= j, logical row i being Y row rowof(i); lane (rg, c) owns
// logical rows i = rg (mod 4) of column c, and lane (0, c) the head row j.
template <typename RowOf>
__device__ void wave_apply_q16(const double* V, int ldv, int rows, int n, const double* vh, const double* beta,
                               double* Y, int ldy, int cols, RowOf rowof) {
  const int l = threadIdx.x & 63, rg = l >> 4, c16 = l & 15;
  for (int c0 = 0; c0 < cols; c0 += 16) {
    const int c = c0 + c16;
    for (int j = n - 1; j >= 0; --j) {
      const double b = beta[j];
      if (b == 0.0) continue;
      const int i0 = j + 1 + ((rg - ((j + 1) & 3)) & 3);
      double p = 0;
      const double* vj = V + (long long)ldv * j;
      double* yc = Y + (long long)ldy * (c < cols ? c : 0);
      if (c < cols) {
        if (rg == 0) p = vh[j] * yc[rowof(j)];
        double p1 = 0, p2 = 0, p3 = 0;
        int i = i0;
        for (; i + 12 < rows; i += 16) {
          p += vj[i] * yc[rowof(i)];
          p1 += vj[i + 4] * yc[rowof(i + 4)];
          p2 += vj[i + 8] * yc[rowof(i + 8)];
          p3 += vj[i + 12] * yc[rowof(i + 12)];
        }
        for (; i < rows; i += 4) p += vj[i] * yc[rowof(i)];
        p = (p + p1) + (p2 + p3);
      }
      p += __shfl_xor(p, 16);
      p += __shfl_xor(p, 32);
      if (c < cols) {
        const double f = b * p;
        if (rg == 0) yc[rowof(j)] -= f * vh[j];
        int i = i0;
        for (; i + 12 < rows; i += 16) {
          const double a0 = vj[i], a1 = vj[i + 4], a2 = vj[i + 8], a3 = vj[i + 12];
          const double y0 = yc[rowof(i)], y1 = yc[rowof(i + 4)], y2 = yc[rowof(i + 8)], y3 = yc[rowof(i + 12)];
          yc[rowof(i)] = y0 - f * a0;
          yc[rowof(i + 4)] = y1 - f * a1;
          yc[rowof(i + 8)] = y2 - f * a2;
          yc[rowof(i + 12)] = y3 - f * a3;
        }
        for (; i < rows; i += 4) yc[rowof(i)] -= f * vj[i];
      }
      // row j - 1, the next head, belongs to another lane
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// One-sided Jacobi SVD of A (M x N, column-major, M >= N) in place:
// A <- U diag(s), V (N x N column-major) accumulated unless V is null.  A pair is rotated while
// |<a_p, a_q>| > tol |a_p| |a_q| with tol = 8 eps sqrt(M), the rounding error
// of the inner product itself, and while both columns are above the rounding
// level of A (|a|^2 > eps^2 ||A||_F^2): a column at that level is zero to
// working precision, and rotating it against its rounding noise would keep
// the sweeps going.  The pairs of a step are disjoint and each group of G
// threads reads and writes only its own rows of its pair, so a step needs one
// barrier (before the next step's pairing).  Returns the number of sweeps.
__device__ int jacobi_svd(double* A, int M, int N, double* V, double* rot, double* red) {
  const int tid = threadIdx.x;
  if (V)
    for (int i = tid; i < N * N; i += blockDim.x) V[i] = (i % N == i / N) ? 1.0 : 0.0;
  double fro = 0;
  for (int i = tid; i < M * N; i += blockDim.x) fro += A[i] * A[i];
  fro = block_sum_d(fro, red);
  if (N == 1) return 0;
  const int ne = N + (N & 1);
  const double tol = 8 * 2.220446049250313e-16 * sqrt((double)M);
  const double floor2 = 2.220446049250313e-16 * 2.220446049250313e-16 * fro;
  int sweep = 0;
  for (; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int st = 0; st < ne - 1; ++st) {
      // each pair handled by a group of G threads
      // (the largest power of two <= 64 with G npair <= blockDim.x: every pair
      // has a group for N <= 2 blockDim.x)
      const int npair = ne / 2;
      int G = 64;
      while (G > 1 && G * npair > (int)blockDim.x) G >>= 1;
      const int pi = tid / G, gl = tid % G;
      double alpha = 0, beta = 0, gamma = 0;
      int p = 0, q = 0;
      const bool active = pi < npair;
      if (active) {
        rr_pair(ne, st, pi, p, q);
        if (q < N) {
          for (int i = gl; i < M; i += G) {
            const double ap = A[i + (long long)M * p], aq = A[i + (long long)M * q];
            alpha += ap * ap;
            beta += aq * aq;
            gamma += ap * aq;
          }
        }
      }
      for (int m = G / 2; m > 0; m >>= 1) {
        alpha += __shfl_xor(alpha, m);
        beta += __shfl_xor(beta, m);
        gamma += __shfl_xor(gamma, m);
      }
      double c = 1, s = 0;
      if (active && q < N && alpha > floor2 && beta > floor2 && fabs(gamma) > tol * sqrt(alpha * beta)) {
        const double zeta = (beta - alpha) / (2 * gamma);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        c = 1 / sqrt(1 + t * t);
        s = c * t;
        rotated = 1;
      }
      if (active && q < N && s != 0.0) {
        for (int i = gl; i < M; i += G) {
          const double ap = A[i + (long long)M * p], aq = A[i + (long long)M * q];
          A[i + (long long)M * p] = c * ap - s * aq;
          A[i + (long long)M * q] = s * ap + c * aq;
        }
        if (V)
          for (int i = gl; i < N; i += G) {
            const double vp = V[i + N * p], vq = V[i + N * q];
            V[i + N * p] = c * vp - s * vq;
            V[i + N * q] = s * vp + c * vq;
          }
      }
      __syncthreads();
    }
    // a barrier that also tells every thread whether any pair rotated (no
    // shared flag to reset, so no thread can see the next sweep's reset)
    if (!__syncthreads_or(rotated)) break;
  }
  return sweep + 1;
}

// Householder QR with column pivoting of A (M x N column-major in LDS,
// N <= M <= 32) by one wave with A in registers, keeping R only: on return A's
// top N x N upper triangle holds R with A P = Q R, and perm[j] is the original
// column in position j.  Lane (c, h) = (l % 32, l / 32) holds rows
// [16 h, 16 h + 16) of column c, as in wave_jacobi_svd32; a step finds the
// remaining column of largest norm (ties: the smaller index), swaps it into
// place and broadcasts the reflector by lane shuffles.  The rows below R hold
// the reflector tails, which nothing reads.
__device__ void wave_qrcp32(double* A, int M, int N, int* perm) {
  const int l = threadIdx.x & 63, c = l & 31, h = l >> 5;
  double a[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = (c < N && 16 * h + i < M) ? A[16 * h + i + M * c] : 0.0;
  int pm = c;  // original column held by this lane's column position
  for (int j = 0; j < N; ++j) {
    // squared norms over rows >= j of the columns c >= j
    double part = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (16 * h + i >= j) part += a[i] * a[i];
    part += __shfl_xor(part, 32);
    double best = (c >= j && c < N) ? part : -1.0;
    int bi = c;
    for (int m = 16; m > 0; m >>= 1) {
      const double ob = __shfl_xor(best, m);
      const int oi = __shfl_xor(bi, m);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    const int p = bi;  // the same in every lane
    // swap columns j and p
    const int src = (c == j ? p : (c == p ? j : c)) + 32 * h;
#pragma unroll
    for (int i = 0; i < 16; ++i) a[i] = __shfl(a[i], src);
    pm = __shfl(pm, src);
    // the reflector of column j (rows >= j): v = [x_jj - alpha; x_{j+1..}]
    double vj[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) vj[i] = __shfl(a[i], j + 32 * h);
    double xloc = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (16 * h + i == j) xloc = vj[i];
    const double xjj = __shfl(xloc, c + 32 * (j >> 4));  // row j lives in half j / 16
    const double nrm = sqrt(best);
    const double alpha = xjj >= 0 ? -nrm : nrm;
    const double v0 = xjj - alpha;
    const double vv = best - xjj * xjj + v0 * v0;
    const double beta = (nrm == 0.0 || vv == 0.0) ? 0.0 : 2.0 / vv;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int row = 16 * h + i;
      vj[i] = row > j ? vj[i] : (row == j ? v0 : 0.0);
    }
    double dot = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) dot += vj[i] * a[i];
    dot += __shfl_xor(dot, 32);
    if (c > j && c < N && beta != 0.0) {
      const double f = beta * dot;
#pragma unroll
      for (int i = 0; i < 16; ++i) a[i] -= f * vj[i];
    }
    if (c == j && beta != 0.0) {
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (16 * h + i == j) a[i] = alpha;
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (c < N && 16 * h + i < M) A[16 * h + i + M * c] = a[i];
  if (h == 0 && c < N) perm[c] = pm;
}

// One-sided Jacobi SVD of A (M x N column-major in LDS, N <= M <= 32) by one
// wave with both matrices in registers: lane (c, h) = (l % 32, l / 32) holds
// rows [16 h, 16 h + 16) of column c of A and of V.  A step fetches the
// partner column of the round-robin schedule (rr_pair's) with lane shuffles,
// joins the two halves' sums with one cross-half shuffle, and both lanes of a
// pair derive the same rotation; the rotation threshold, the rounding floor and
// the stopping rule are those of jacobi_svd.  A step needs no LDS and no
// barrier (jacobi_svd's steps are LDS round trips with a workgroup barrier
// each, ~1.3 us).  Writes A <- U diag(s) and, with WANT_V, V (N x N) back;
// returns the number of sweeps.  Every lane of the wave must call it.
template <bool WANT_V>
__device__ int wave_jacobi_svd32(double* A, int M, int N, double* V) {
  const int l = threadIdx.x & 63, c = l & 31, h = l >> 5;
  double a[16], v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = 16 * h + i;
    a[i] = (c < N && row < M) ? A[row + M * c] : 0.0;
    v[i] = (c < N && row == c) ? 1.0 : 0.0;
  }
  double fro = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) fro += a[i] * a[i];
  fro = wave_sum_d(fro);
  const int ne = N + (N & 1);
  const double tol = 8 * 2.220446049250313e-16 * sqrt((double)M);
  const double floor2 = 2.220446049250313e-16 * 2.220446049250313e-16 * fro;
  int sweep = 0;
  for (; sweep < 60 && N > 1; ++sweep) {
    bool rotated = false;
    for (int st = 0; st < ne - 1; ++st) {
      // partner of column c in step st (the inverse of rr_pair's placement)
      int q = c;
      if (c < ne) {
        int t = c - 1 - st;
        if (t < 0) t += ne - 1;
        const int pos = c == 0 ? 0 : 1 + t;
        const int pp = ne - 1 - pos;
        int u = pp - 1 + st;
        if (u >= ne - 1) u -= ne - 1;
        q = pp == 0 ? 0 : 1 + u;
      }
      const int src = q + 32 * h;
      double b[16];
      double own = 0, oth = 0, cr = 0;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        b[i] = __shfl(a[i], src);
        own += a[i] * a[i];
        oth += b[i] * b[i];
        cr += a[i] * b[i];
      }
      own += __shfl_xor(own, 32);
      oth += __shfl_xor(oth, 32);
      cr += __shfl_xor(cr, 32);
      const bool isp = c < q;  // this lane holds the pair's first column
      const double alpha = isp ? own : oth, beta = isp ? oth : own;
      // first column: c a_p - s a_q; second: s a_p + c a_q (the partner is the
      // other); a lane that does not rotate keeps c = 1, s = 0 (exact)
      double cs = 1.0, sn = 0.0;
      if (c < N && q < N && c != q && alpha > floor2 && beta > floor2 && fabs(cr) > tol * sqrt(alpha * beta)) {
        const double zeta = (beta - alpha) / (2 * cr);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        cs = 1 / sqrt(1 + t * t);
        sn = cs * t;
        rotated = true;
      }
      const double sp = isp ? -sn : sn;
#pragma unroll
      for (int i = 0; i < 16; ++i) a[i] = cs * a[i] + sp * b[i];
      if (WANT_V) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const double wi = __shfl(v[i], src);
          v[i] = cs * v[i] + sp * wi;
        }
      }
    }
    if (__ballot(rotated) == 0) break;
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = 16 * h + i;
    if (c < N && row < M) A[row + M * c] = a[i];
    if (WANT_V && c < N && row < N) V[row + N * c] = v[i];
  }
  return sweep + 1;
}

// ---- Gram-based QR of a side, global-memory variant ----------------------
// The QR of a tall side enters simple_update_bp only through R (in Theta) and
// Q (in the rebuild), and any factorisation X = Q R with orthonormal Q gives
// the same psi' (R changes by an orthogonal factor that Theta's SVD absorbs).
// Here: G = X^T X = V diag(l) V^T (one-sided Jacobi of the symmetric G),
// directions with l <= n eps l_max dropped (rank r); X1 = X T with
// T = V_r diag(l_r)^{-1/2}; G1 = X1^T X1 = R1^T R1 (Cholesky, G1 ~ I); then
// R = R1 diag(l_r)^{1/2} V_r^T (r x n) and Q = X Tq, Tq = T R1^{-1}: the
// second Gram removes the first one's loss of orthogonality (the scheme of
// CholeskyQR2).  Directions of X below ~sqrt(n eps) of its largest singular
// value are dropped, as rank-deficient ones.  Every step streams X through
// the workgroup from global memory a fixed number of times, where the
// column-by-column Householder QR reads it once per column.

// G = X^T X (n x n, ld n, in LDS) of X (E x n column-major, global), the rows
// staged through `tile` (tile_cap doubles of LDS) in chunks, the products on
// the FP64 matrix cores: wave w owns the 16-row block w of G and its four
// 16 x 16 tiles (n <= 64), v_mfma_f64_16x16x4f64 with A[i][k] = X[e + k][16 w + i]
// and B[k][j] = X[e + k][16 bj + j] for four rows e.. e + 3 per step (rows past
// the chunk and columns past n read as zero); D[4 v + l / 16][l % 16] is value
// v of lane l
__device__ void gram_rows(const double* X, long long E, int n, double* G, double* tile, int tile_cap) {
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, li = l & 15, lk = l >> 4;
  const int ld = tile_cap / n;  // rows of a chunk + 1 (odd pitch against bank conflicts)
  const int cr = (ld & 1) ? ld - 1 : ld - 2;
  const int pitch = cr + 1;
  f64x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f64x4{0, 0, 0, 0};
  const int ri = 16 * wv + li;  // this lane's row of G (A operand column of X)
  for (long long e0 = 0; e0 < E; e0 += cr) {
    const int rows = (int)((E - e0) < cr ? (E - e0) : cr);
    __syncthreads();  // the previous chunk is consumed
    for (int idx = tid; idx < rows * n; idx += blockDim.x) {
      const int c = idx / rows, rr = idx - c * rows;
      tile[c * pitch + rr] = X[e0 + rr + E * c];
    }
    __syncthreads();
    if (16 * wv < n)
      for (int k0 = 0; k0 < rows; k0 += 4) {
        const int rr = k0 + lk;
        const bool rin = rr < rows;
        const double av = (rin && ri < n) ? tile[ri * pitch + rr] : 0.0;
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
          const int cj = 16 * bj + li;
          const double bv = (rin && cj < n) ? tile[cj * pitch + rr] : 0.0;
          acc[bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[bj], 0, 0, 0);
        }
      }
  }
  __syncthreads();
#pragma unroll
  for (int bj = 0; bj < 4; ++bj)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * wv + 4 * v + lk, col = 16 * bj + li;
      if (row < n && col < n) G[row + n * col] = acc[bj][v];
    }
  __syncthreads();
}

// out[e + E j] = sum_c X[e + E c] W[c + n j] for j < m (n <= 64), W (n x m,
// ld n) in LDS, on the FP64 matrix cores: wave w takes the 16-row blocks
// w, w + 4, ..; a block's A operands (A[i][k] = X[e0 + i][4 s + k], lane
// i + 16 k, one register per step s) are loaded once and reused for every
// 16-column tile of the output (B[k][j] = W[4 s + k][16 t + j]); rows, columns
// and steps past E, m and n read as zero.  D[4 v + l / 16][l % 16] is value v
// of lane l.
__device__ void rows_times(const double* X, long long E, int n, const double* W, int m, double* out) {
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, li = l & 15, lk = l >> 4;
  const int nsteps = (n + 3) >> 2;  // <= 16
  for (long long e0 = 16LL * wv; e0 < E; e0 += 64) {
    double av[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int c = 4 * s + lk;
      av[s] = (s < nsteps && c < n && e0 + li < E) ? X[e0 + li + E * c] : 0.0;
    }
    for (int t0 = 0; t0 < m; t0 += 16) {
      f64x4 acc = {0, 0, 0, 0};
      const int j = t0 + li;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        if (s < nsteps) {
          const int c = 4 * s + lk;
          const double bv = (c < n && j < m) ? W[c + n * j] : 0.0;
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long e = e0 + 4 * v + lk;
        if (e < E && j < m) out[e + E * j] = acc[v];
      }
    }
  }
  __syncthreads();
}

// The factorisation above for one side.  X (E x n, global) is left unchanged;
// Xw (E x n, global) is work; lds: 4 n^2 + 4 n doubles of LDS (>= 4 AP_N_BIG^2
// + ...); R (r x n, ld r) and Tq (n x r, ld n) to global, lam: n doubles of
// LDS.  Returns r.
__device__ int gram_qr(const double* X, long long E, int n, double* Xw, double* lds, int lds_cap, double* lam,
                       int* keep_idx, double* R, double* Tq, double* rot, double* red) {
  const int tid = threadIdx.x;
  double* G = lds;             // n x n: G, then U diag(l); then G1 / R1 (r x r)
  double* V = G + n * n;       // n x n
  double* T = V + n * n;       // n x r
  double* work = T + n * n;    // Gram staging
  const int wcap = lds_cap - 3 * n * n;
  __shared__ int s_rank;
  gram_rows(X, E, n, G, T, lds_cap - 2 * n * n);
  if (n <= 32) {
    // one wave, QR-preconditioned (as for Theta): G P = Q R, Jacobi on R^T ->
    // U_x diag(l); G's eigenvectors are P U_x, its eigenvalues the column norms
    __shared__ int gperm[32];
    double* Xr = T;  // free until T is built
    if (tid < 64) wave_qrcp32(G, n, n, gperm);
    __syncthreads();
    for (int i = tid; i < n * n; i += blockDim.x) {
      const int row = i % n, col = i / n;
      Xr[i] = col <= row ? G[col + n * row] : 0.0;
    }
    __syncthreads();
    if (tid < 64) wave_jacobi_svd32<false>(Xr, n, n, nullptr);
    __syncthreads();
    for (int j = tid; j < n; j += blockDim.x) {
      double t = 0;
      for (int i = 0; i < n; ++i) t += Xr[i + n * j] * Xr[i + n * j];
      lam[j] = sqrt(t);
    }
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += blockDim.x) {
      const int i = idx % n, j = idx / n;
      V[gperm[i] + n * j] = lam[j] > 0 ? Xr[i + n * j] / lam[j] : 0.0;
    }
    __syncthreads();
  } else {
    jacobi_svd(G, n, n, V, rot, red);
    // eigenvalues of the PSD G: the column norms of U diag(l)
    for (int j = tid; j < n; j += blockDim.x) {
      double t = 0;
      for (int i = 0; i < n; ++i) t += G[i + n * j] * G[i + n * j];
      lam[j] = sqrt(t);
    }
    __syncthreads();
  }
  if (tid == 0) {
    double lmax = 0;
    for (int j = 0; j < n; ++j) lmax = lam[j] > lmax ? lam[j] : lmax;
    int r = 0;
    for (int j = 0; j < n; ++j)
      if (lam[j] > n * 2.220446049250313e-16 * lmax) keep_idx[r++] = j;
    if (r == 0) keep_idx[r++] = 0;  // X = 0: one zero direction, R = 0
    s_rank = r;
  }
  __syncthreads();
  const int r = s_rank;
  for (int idx = tid; idx < n * r; idx += blockDim.x) {
    const int c = idx % n, k = idx / n;
    const double l = lam[keep_idx[k]];
    T[idx] = l > 0 ? V[c + n * keep_idx[k]] / sqrt(l) : 0.0;
  }
  __syncthreads();
  rows_times(X, E, n, T, r, Xw);  // X1 = X T
  gram_rows(Xw, E, r, G, work, wcap);  // G1 = X1^T X1 (r x r, ld r)
  // Cholesky G1 = R1^T R1 in place (upper triangle, ld r); a non-positive pivot
  // (a zero column of X1) gives a unit diagonal and a zero row
  for (int k = 0; k < r; ++k) {
    const double pkk = G[k + r * k];
    const double piv = pkk > 0 ? sqrt(pkk) : 0.0;
    __syncthreads();
    for (int j = k + tid; j < r; j += blockDim.x)
      G[k + r * j] = j == k ? (piv > 0 ? piv : 1.0) : (piv > 0 ? G[k + r * j] / piv : 0.0);
    __syncthreads();
    const int m = r - k - 1;
    for (int idx = tid; idx < m * m; idx += blockDim.x) {
      const int i = k + 1 + idx % m, j = k + 1 + idx / m;
      if (i <= j) G[i + r * j] -= G[k + r * i] * G[k + r * j];
    }
    __syncthreads();
  }
  // R = R1 diag(l_r)^{1/2} V_r^T
  for (int idx = tid; idx < r * n; idx += blockDim.x) {
    const int i = idx % r, c = idx / r;
    double t = 0;
    for (int k = i; k < r; ++k) t += G[i + r * k] * sqrt(lam[keep_idx[k]]) * V[c + n * keep_idx[k]];
    R[i + r * c] = t;
  }
  // Tq = T R1^{-1}: forward substitution along each row of T
  for (int c = tid; c < n; c += blockDim.x)
    for (int j = 0; j < r; ++j) {
      double t = T[c + n * j];
      for (int k = 0; k < j; ++k) t -= Tq[c + n * k] * G[k + r * j];
      Tq[c + n * j] = t / G[j + r * j];
    }
  __syncthreads();
  return r;
}

// BIG = false: the side matrices (E n <= AP_XY) and every small matrix in LDS,
// the Householder reflectors of each side saved to `work` (2 AP_SIDE doubles
// per gate).  BIG = true: the same steps with every matrix in a global
// workspace of prm.big_stride doubles per gate at `work` (layout below), for
// external bonds up to AP_ENV_BIG, d chi_b <= AP_N_BIG and Theta up to
// AP_TH_BIG square; LDS holds the geometry, the vectors and the environment
// eigenproblems.
template <bool BIG>
__global__ __launch_bounds__(256) void k_apply_gates(const ApplyItem* __restrict__ items,
                                                     const VDesc* __restrict__ vds,
                                                     double* __restrict__ psi,
                                                     const double* __restrict__ msg,
                                                     const double* __restrict__ gates, ApplyParams prm,
                                                     double* __restrict__ work,
                                                     double* __restrict__ truncerr,
                                                     double* __restrict__ svals, int* __restrict__ kept,
                                                     int* __restrict__ status) {
  constexpr int TH = BIG ? AP_TH_BIG : AP_TH;           // rows/cols of Theta
  constexpr int ENV = BIG ? AP_ENV_BIG : AP_ENV;        // external bond dimension
  constexpr int RS = ENV * ENV;                         // one environment root
  static_assert(4 * ENVB_WORK <= AP_TH_BIG * AP_TH_BIG, "environment work must fit the Theta area");
  static_assert(2 * AP_F_BIG >= AP_TH_BIG * AP_TH_BIG, "a copy of Theta must fit the R factors' area");
  extern __shared__ __align__(16) double sm[];
  double *X, *Y, *Ru, *Rv, *roots, *small, *A, *VV, *scr_u, *scr_v;
  double *tq_u = nullptr, *tq_v = nullptr;  // BIG: Q = X Tq of each side
  if constexpr (BIG) {
    // per gate in global memory: QR of side u (E_u n_u), of side v, two work
    // matrices (E max(n, d keep)), the R / factor matrices, V of the SVD
    // (when it does not fit in LDS beside Theta), the roots; in LDS the small
    // vectors, the matrix of a mode product and Theta / U diag(s)
    double* w = work + (long long)blockIdx.x * prm.big_stride;
    scr_u = w;
    scr_v = scr_u + prm.big_q;
    X = scr_v + prm.big_q;
    Y = X + prm.big_x;
    Ru = Y + prm.big_x;
    Rv = Ru + AP_F_BIG;
    VV = Rv + AP_F_BIG;
    roots = VV + AP_TH_BIG * AP_TH_BIG;
    tq_u = roots + 2 * AP_MAXEXT * 2 * RS;
    tq_v = tq_u + AP_N_BIG * AP_N_BIG;
    small = sm;
    A = sm + AP_BIG_SMALL + AP_ENV_BIG * AP_ENV_BIG;
  } else {
    X = sm;                       // AP_XY: QR, then SVD workspace
    Y = X + AP_XY;                // AP_XY: mode-product partner, rebuild
    Ru = Y + AP_XY;               // 1024: R_u (r x n), then the left factor
    Rv = Ru + 1024;               // 1024: R_v, then the right factor
    roots = Rv + 1024;            // 2 sides x AP_MAXEXT x (sqrt, inv sqrt) x AP_ENV^2
    small = roots + 2 * AP_MAXEXT * 2 * RS;
    A = X;                        // Theta, then U diag(s)
    VV = X + 1024;                // V of the SVD
    scr_u = work + (long long)blockIdx.x * 2 * AP_SIDE;
    scr_v = scr_u + AP_SIDE;
  }
  double* rot = small;
  double* vh = rot + 4 * 32;
  double* beta = vh + AP_N_BIG;
  double* vhv = beta + AP_N_BIG;
  double* betav = vhv + AP_N_BIG;
  double* svec = betav + AP_N_BIG;
  double* red = svec + TH;
  double* ms = small + AP_BIG_SMALL;  // BIG: the matrix of a mode product
  double* envw = A;                   // BIG: ENVB_WORK doubles per wave, phase 1 only
  __shared__ int sh_keep;
  __shared__ int order[TH];
  __shared__ int gkeep[BIG ? AP_N_BIG : 1];  // BIG: the directions the Gram QR keeps
  __shared__ int qgemm[2];                    // LDS kernel: side rebuilt as X R^{-1} F

  const ApplyItem it = items[blockIdx.x];
  const int tid = threadIdx.x;
  // the factorisation runs on the bond as it was (columns b < chib_in); a bond
  // grown for the kept rank only widens the layout the result is written to
  // both sides' geometry in LDS: its arrays are indexed at run time, which
  // would put a per-thread copy in scratch memory
  __shared__ SideGeom sgeo[2];
  if (tid < 2) {
    SideGeom g = side_geom(vds[tid == 0 ? it.u : it.v], tid == 0 ? it.ju : it.jv);
    g.chib = it.chib_in;
    g.n = g.d * g.chib;
    g.r = g.E < g.n ? g.E : g.n;
    sgeo[tid] = g;
  }
  __syncthreads();
  const SideGeom& gu = sgeo[0];
  const SideGeom& gv = sgeo[1];
  phase_mark(prm, 0);

  // 1. environment roots (sqrt and inverse sqrt), both sides, one environment per wave
  {
    const int wv = tid >> 6;
    for (int q = wv; q < gu.next + gv.next; q += (int)blockDim.x / 64) {
      const int side = q < gu.next ? 0 : 1;
      const int pa = side == 0 ? q : q - gu.next;
      const SideGeom& g = side == 0 ? gu : gv;
      const VDesc& V = vds[side == 0 ? it.u : it.v];
      const int a = g.ext[pa];
      double* S = roots + ((side * AP_MAXEXT + pa) * 2) * RS;
      if constexpr (BIG)
        env_roots_big(msg + V.in_msg[a], g.dims[a], prm.eig_cutoff, S, S + RS, envw + wv * ENVB_WORK);
      else
        env_roots_wave(msg + V.in_msg[a], g.dims[a], prm.eig_cutoff, S, S + RS);
    }
    __syncthreads();
  }
  phase_mark(prm, 1);
  // 2. X = psi with sqrt(env) on the external bonds, QR; reflectors to scratch
  for (int side = 0; side < 2; ++side) {
    const SideGeom& g = side == 0 ? gu : gv;
    for (int idx = tid; idx < g.E * g.n; idx += blockDim.x) {
      const int c = (int)g.fE.div(idx), e = idx - c * g.E;
      const int b = (int)g.fd.div(c), s = c - b * g.d;
      X[idx] = psi[g.psi_off + nat_off(g, s, e, b)];
    }
    __syncthreads();
    if (side == 0) phase_mark(prm, 10);
    double* cur = X;
    double* oth = Y;
    for (int pa = 0; pa < g.next; ++pa) {
      const double* S = roots + ((side * AP_MAXEXT + pa) * 2) * RS;
      if constexpr (BIG)
        ext_mode_product_fibre<ENV>(g, pa, cur, oth, g.n, S, false, ms);
      else
        ext_mode_product<ENV>(g, pa, cur, oth, g.n, S, false);
      double* t = cur;
      cur = oth;
      oth = t;
    }
    if (cur != X) {
      for (int i = tid; i < g.E * g.n; i += blockDim.x) X[i] = cur[i];
      __syncthreads();
    }
    double* vhs = side == 0 ? vh : vhv;
    double* bts = side == 0 ? beta : betav;
    double* R = side == 0 ? Ru : Rv;
    double* scr = side == 0 ? scr_u : scr_v;
    if (!BIG) {
      // X itself, for a rebuild by Q = X R^{-1} (kept if the side qualifies)
      for (int i = tid; i < g.E * g.n; i += blockDim.x) scr[i] = X[i];
      __syncthreads();
    }
    if (side == 0) phase_mark(prm, 11);
    if (!BIG && tsqr_ok(g)) {
      // the stacked R factors S (4n x n) in Y, which is free after the copy to X
      const int n4 = 4 * g.n, wv = tid >> 6;
      double* S = Y;
      if (wv < 4) {
        const int r0 = wv * g.E / 4, r1 = (wv + 1) * g.E / 4;
        wave_qr16(X + r0, g.E, r1 - r0, g.n, vhs + wv * g.n, bts + wv * g.n);
        for (int idx = (tid & 63); idx < g.n * g.n; idx += 64) {
          const int i = idx % g.n, c = idx / g.n;
          S[wv * g.n + i + n4 * c] = (i <= c) ? X[r0 + i + (long long)g.E * c] : 0.0;
        }
      }
      __syncthreads();
      double* qx = scr + AP_XY;  // S's reflectors, then their heads and betas
      if (wv == 0) wave_qr16(S, n4, n4, g.n, qx + n4 * g.n, qx + n4 * g.n + 16);
      __syncthreads();
      for (int idx = tid; idx < g.r * g.n; idx += blockDim.x) {
        const int i = idx % g.r, c = idx / g.r;
        R[i + g.r * c] = (i <= c) ? S[i + n4 * c] : 0.0;
      }
      for (int i = tid; i < n4 * g.n; i += blockDim.x) qx[i] = S[i];
    } else {
      if constexpr (BIG) {
        // the Gram QR: X unchanged, Y as work, the rank r replaces min(E, n)
        const int r = gram_qr(X, g.E, g.n, Y, A, AP_TH_BIG * AP_TH_BIG, svec, gkeep, R, side == 0 ? tq_u : tq_v,
                              rot, red);
        if (tid == 0) sgeo[side].r = r;
        __syncthreads();
      } else {
        householder_qr(X, g.E, g.n, vhs, bts, red);
        for (int idx = tid; idx < g.r * g.n; idx += blockDim.x) {
          const int i = idx % g.r, c = idx / g.r;
          R[i + g.r * c] = (i <= c) ? X[i + (long long)g.E * c] : 0.0;  // R[i][c], column-major (ld r)
        }
      }
    }
    if (side == 0) phase_mark(prm, 12);
    if constexpr (BIG) {
      for (int i = tid; i < g.E * g.n; i += blockDim.x) scr[i] = X[i];
    } else {
      // a tall side whose R has no diagonal entry below 1e-5 of its largest
      // keeps X and R: the rebuild's Q [F; 0] is then X (R^{-1} F), one small
      // product, instead of applying the n reflectors (latency-bound), with an
      // orthogonality error ~eps cond(X); other sides keep their reflectors
      __syncthreads();  // R is complete
      if (tid == 0) {
        bool ok = g.E >= g.n;
        double dmax = 0, dmin = INFINITY;
        for (int j = 0; j < g.n && ok; ++j) {
          const double dj = fabs(R[j + g.r * j]);
          dmax = dj > dmax ? dj : dmax;
          dmin = dj < dmin ? dj : dmin;
        }
        qgemm[side] = ok && dmax > 0 && isfinite(dmax) && dmin >= 1e-5 * dmax;
      }
      __syncthreads();
      if (qgemm[side]) {
        for (int i = tid; i < g.n * g.n; i += blockDim.x) scr[AP_XY + i] = R[i];  // ld n (r = n)
      } else {
        for (int i = tid; i < g.E * g.n; i += blockDim.x) scr[i] = X[i];  // the reflectors
      }
    }
    __syncthreads();
    if (side == 0) phase_mark(prm, 13);
  }
  phase_mark(prm, 2);
  // 3. Theta[(ru, su'), (rv, sv')] = sum G[(su',sv'),(su,sv)] Ru[ru, su + du b] Rv[rv, sv + dv b],
  //    stored directly in the taller orientation (X: M x N column-major), V in X + 1024
  const int du = gu.d, dv = gv.d, chib = gu.chib;
  const int Mr = gu.r * du, Nc = gv.r * dv;
  const int dd = du * dv;
  const double* G = gates + it.gate;
  const bool tr = Mr < Nc;
  const int M = tr ? Nc : Mr, N = tr ? Mr : Nc;
  if constexpr (BIG)
    if (M * N + N * N <= AP_TH_BIG * AP_TH_BIG) VV = A + M * N;  // else in the global workspace
  for (int idx = tid; idx < Mr * Nc; idx += blockDim.x) {
    const int row = idx % Mr, col = idx / Mr;
    const int ru = row % gu.r, sup = row / gu.r;
    const int rv = col % gv.r, svp = col / gv.r;
    double acc = 0;
    for (int su = 0; su < du; ++su)
      for (int sv = 0; sv < dv; ++sv) {
        const double gval = G[(sup + du * svp) + dd * (su + du * sv)];
        if (gval == 0.0) continue;
        double rr = 0;
        for (int b = 0; b < chib; ++b) rr += Ru[ru + gu.r * (su + du * b)] * Rv[rv + gv.r * (sv + dv * b)];
        acc += gval * rr;
      }
    if (tr) A[col + Nc * row] = acc;
    else A[idx] = acc;
  }
  __syncthreads();
  phase_mark(prm, 3);
  // 4. SVD
  // Theta up to 32 x 32 (QR-preconditioned one-sided Jacobi, one wave):
  // Theta P = Q R by a pivoted Householder QR, then Jacobi on X = R^T in
  // registers, X <- U_x diag(s) (R^T = U_x S V_x^T, so Theta = (Q V_x) S (P U_x)^T):
  // V_Theta = P U_x, and the kept U_Theta s = Theta V_Theta from a copy of
  // Theta, so neither Q nor V_x is formed.  The pivoting grades X's columns and
  // halves the sweeps (7 instead of 13-14 on a config-2 Theta).  Larger Theta
  // (global-memory variant): the workgroup's Jacobi on Theta, V accumulated.
  int nsweeps = 0;
  const bool pre = M <= 32;
  double* Th0 = A + 2048;  // pre: a copy of Theta, then X = R^T at Th0 + 1024
  double* Xr = Th0 + 1024;
  __shared__ int perm[TH];
  if (pre) {
    for (int i = tid; i < M * N; i += blockDim.x) Th0[i] = A[i];
    __syncthreads();
    if (tid < 64) wave_qrcp32(A, M, N, perm);
    __syncthreads();
    for (int i = tid; i < N * N; i += blockDim.x) {
      const int row = i % N, col = i / N;
      Xr[i] = col <= row ? A[col + M * row] : 0.0;  // X[row][col] = R[col][row]
    }
    __syncthreads();
    if (tid < 64) nsweeps = wave_jacobi_svd32<false>(Xr, N, N, nullptr);
    __syncthreads();
  } else {
    if constexpr (BIG) {
      // Theta > 32 x 32 (global-memory variant): V of the kept columns from a
      // copy of Theta afterwards, as for the preconditioned path, which spares
      // the rotation of V (N x N, in global memory at N = 128) in every step;
      // the copy goes where R_u and R_v were (free once Theta is built)
      for (int i = tid; i < M * N; i += blockDim.x) Ru[i] = A[i];
      __syncthreads();
      nsweeps = jacobi_svd(A, M, N, nullptr, rot, red);
    } else {
      nsweeps = jacobi_svd(A, M, N, VV, rot, red);
    }
  }
  if (prm.ts && tid == 0) {
    prm.ts[blockIdx.x * 16 + 8] = (unsigned long long)nsweeps;
    prm.ts[blockIdx.x * 16 + 9] = (unsigned long long)((!BIG && tsqr_ok(gu)) + (!BIG && tsqr_ok(gv)));
  }
  for (int j = tid; j < N; j += blockDim.x) {
    double s = 0;
    if (pre)
      for (int i = 0; i < N; ++i) s += Xr[i + N * j] * Xr[i + N * j];
    else
      for (int i = 0; i < M; ++i) s += A[i + M * j] * A[i + M * j];
    svec[j] = sqrt(s);
  }
  __syncthreads();
  if (tid == 0) {
    for (int i = 0; i < N; ++i) order[i] = i;
    for (int i = 0; i < N; ++i)
      for (int j = i + 1; j < N; ++j)
        if (svec[order[j]] > svec[order[i]]) {
          int t = order[i];
          order[i] = order[j];
          order[j] = t;
        }
    double tot = 0;
    for (int i = 0; i < N; ++i) tot += svec[i] * svec[i];
    int keep = N;
    if (prm.maxdim > 0 && keep > prm.maxdim) keep = prm.maxdim;
    double disc = 0;
    for (int i = keep; i < N; ++i) disc += svec[order[i]] * svec[order[i]];
    if (prm.cutoff >= 0 && tot > 0) {
      while (keep > 1 && (disc + svec[order[keep - 1]] * svec[order[keep - 1]]) / tot <= prm.cutoff) {
        disc += svec[order[keep - 1]] * svec[order[keep - 1]];
        --keep;
      }
    }
    int st = 0;
    if (keep > gu.dims[gu.j]) st = 1;  // more than the bond's layout holds
    if (!isfinite(tot)) st = 2;  // non-finite Theta (a NaN or Inf in psi or the environment)
    truncerr[it.slot] = tot > 0 ? disc / tot : (st == 2 ? tot : 0.0);
    if (svals)
      for (int i = 0; i < prm.max_keep; ++i)
        svals[(long long)it.slot * prm.max_keep + i] = i < keep && i < N ? svec[order[i]] : 0.0;
    if (kept) kept[it.slot] = keep;
    status[it.slot] = st;
    sh_keep = st ? -1 : keep;
  }
  __syncthreads();
  phase_mark(prm, 4);
  const int keep = sh_keep;
  // a bond whose kept rank exceeds its dimension, or whose Theta is not
  // finite, is left untouched (the host reports the error); every write to psi
  // comes after this point
  if (keep < 0) return;
  if (BIG && !pre) {
    // the kept right singular vectors: V_o = Theta^T A_o / s_o^2 (zero for s_o = 0)
    for (int idx = tid; idx < N * keep; idx += blockDim.x) {
      const int i = idx % N, o = order[idx / N];
      const double s2 = svec[o] * svec[o];
      double t = 0;
      for (int r = 0; r < M; ++r) t += Ru[r + (long long)M * i] * A[r + (long long)M * o];
      VV[i + N * o] = s2 > 0 ? t / s2 : 0.0;
    }
    __syncthreads();
  }
  if (pre) {
    // the kept singular vectors: V_o = P U_x,o = P X_o / s_o, then
    // A_o = U_o s_o = Theta V_o (a zero singular value gets zero columns, which
    // its sqrt(s) = 0 factor multiplies anyway)
    for (int idx = tid; idx < N * keep; idx += blockDim.x) {
      const int i = idx % N, o = order[idx / N];
      const double so = svec[o];
      VV[perm[i] + N * o] = so > 0 ? Xr[i + N * o] / so : 0.0;
    }
    __syncthreads();
    for (int idx = tid; idx < M * keep; idx += blockDim.x) {
      const int r = idx % M, o = order[idx / M];
      double t = 0;
      for (int i = 0; i < N; ++i) t += Th0[r + M * i] * VV[i + N * o];
      A[r + M * o] = t;
    }
    __syncthreads();
  }
  // left factor Lf (Mr x keep) = U sqrt(S) into Ru, right factor Rf (Nc x keep) = V sqrt(S) into Rv;
  // A = U diag(s) (M x N), V = VV (N x N); if transposed, Theta = VV S U^T
  for (int idx = tid; idx < Mr * keep; idx += blockDim.x) {
    const int i = idx % Mr, j = idx / Mr;
    const int o = order[j];
    const double sv = svec[o];
    Ru[idx] = !tr ? (sv > 0 ? A[i + M * o] / sqrt(sv) : 0.0) : VV[i + N * o] * sqrt(sv);
  }
  for (int idx = tid; idx < Nc * keep; idx += blockDim.x) {
    const int i = idx % Nc, j = idx / Nc;
    const int o = order[j];
    const double sv = svec[o];
    Rv[idx] = !tr ? VV[i + N * o] * sqrt(sv) : (sv > 0 ? A[i + M * o] / sqrt(sv) : 0.0);
  }
  __syncthreads();
  phase_mark(prm, 5);
  // 5. rebuild both sides: Y = [F as r x (d keep); 0], apply Q, then M^{-1/2};
  //    the largest-magnitude entry of every new-bond slice of psi_u' (first in
  //    column-major order) is made positive and the same sign applied to psi_v'
  __shared__ double bond_sign[BIG ? AP_KEEP_BIG : AP_TH];
  for (int side = 0; side < 2; ++side) {
    const SideGeom& g = side == 0 ? gu : gv;
    const double* F = side == 0 ? Ru : Rv;  // rows (r_i, s), r fastest
    const double* scr = side == 0 ? scr_u : scr_v;
    const int cols = g.d * keep;  // column c' = s + d b'
    if constexpr (BIG) {
      // Q [F] = X Tq F: W = Tq F (n x cols) in the Theta area (free once the
      // factors are out), then Y = X W with X (psi with sqrt(env)) from scr
      const double* Tq = side == 0 ? tq_u : tq_v;
      for (int idx = tid; idx < g.n * cols; idx += blockDim.x) {
        const int c = idx % g.n, cc = idx / g.n;
        const int b = cc / g.d, s = cc - b * g.d;
        double t = 0;
        for (int k = 0; k < g.r; ++k) t += Tq[c + g.n * k] * F[(k + g.r * s) + (long long)(g.r * g.d) * b];
        A[idx] = t;
      }
      __syncthreads();
      rows_times(scr, g.E, g.n, A, cols, Y);
    } else if (qgemm[side]) {
      // W = R^{-1} F (n x cols, back substitution per column) in LDS X, Y = X_side W
      const double* Rg = scr + AP_XY;
      for (int cc = tid; cc < cols; cc += blockDim.x) {
        const int b = cc / g.d, s = cc - b * g.d;
        for (int i = g.n - 1; i >= 0; --i) {
          double t = F[(i + g.n * s) + (long long)(g.n * g.d) * b];
          for (int k = i + 1; k < g.n; ++k) t -= Rg[i + g.n * k] * X[k + g.n * cc];
          X[i + g.n * cc] = t / Rg[i + g.n * i];
        }
      }
      __syncthreads();
      rows_times(scr, g.E, g.n, X, cols, Y);
    } else {
      // the side's reflectors, copied back into LDS
      for (int i = tid; i < g.E * g.n; i += blockDim.x) X[i] = scr[i];
      for (int idx = tid; idx < g.E * cols; idx += blockDim.x) {
        const int c = (int)g.fE.div(idx), e = idx - c * g.E;
        const int b = (int)g.fd.div(c), s = c - b * g.d;
        Y[idx] = e < g.r ? F[(e + g.r * s) + (long long)(g.r * g.d) * b] : 0.0;
      }
      __syncthreads();
      if (tsqr_ok(g)) {
        // Q [F; 0] = diag(Q_w) P Q_S [F; 0]: Q_S on the stacked rows (block w's
        // first n rows) by wave 0, then each block's reflectors by its wave
        const int n4 = 4 * g.n, wv = tid >> 6, E = g.E, nn = g.n;
        const double* qx = scr + AP_XY;
        // stacked row i -> block i / nn, row i % nn (i / nn by a multiply-high: i nn < 2^16)
        const unsigned inv = (65536u + nn - 1) / nn;
        if (wv == 0)
          wave_apply_q16(qx, n4, n4, nn, qx + n4 * nn, qx + n4 * nn + 16, Y, E, cols, [=](int i) {
            const int w = (int)(((unsigned)i * inv) >> 16);
            return w * E / 4 + (i - w * nn);
          });
        __syncthreads();
        if (wv < 4) {
          const int r0 = wv * E / 4, r1 = (wv + 1) * E / 4;
          const double* vb = side == 0 ? vh : vhv;
          const double* bb = side == 0 ? beta : betav;
          wave_apply_q16(X + r0, E, r1 - r0, nn, vb + wv * nn, bb + wv * nn, Y + r0, E, cols, [](int i) { return i; });
        }
        __syncthreads();
      } else {
        apply_q(X, g.E, g.r, side == 0 ? vh : vhv, side == 0 ? beta : betav, Y, cols);
      }
    }
    // M^{-1/2}: psi'[.. b ..] = sum_b' Y[.. b' ..] Si[b][b']
    double* cur = Y;
    double* oth = X;
    for (int pa = 0; pa < g.next; ++pa) {
      const double* Si = roots + ((side * AP_MAXEXT + pa) * 2 + 1) * RS;
      if constexpr (BIG)
        ext_mode_product_fibre<ENV>(g, pa, cur, oth, cols, Si, true, ms);
      else
        ext_mode_product<ENV>(g, pa, cur, oth, cols, Si, true);
      double* t = cur;
      cur = oth;
      oth = t;
    }
    if (side == 0) {
      // one wave per new-bond column: the largest |entry| (ties: the smallest
      // natural offset, so the choice does not depend on the scan order)
      const int wv = tid >> 6, l = tid & 63, nw = (int)(blockDim.x >> 6);
      for (int j = wv; j < keep; j += nw) {
        double best = -1;
        int best_off = 0x7fffffff;
        double val = 0;
        for (int idx = l; idx < g.E * g.d; idx += 64) {
          const int s2 = (int)g.fE.div(idx), e = idx - s2 * g.E;
          const double x = cur[e + (long long)g.E * (s2 + g.d * j)];
          const int off = nat_off(g, s2, e, j);
          if (fabs(x) > best || (fabs(x) == best && off < best_off)) {
            best = fabs(x);
            best_off = off;
            val = x;
          }
        }
        for (int m = 32; m > 0; m >>= 1) {
          const double ob = __shfl_xor(best, m), ov = __shfl_xor(val, m);
          const int oo = __shfl_xor(best_off, m);
          if (ob > best || (ob == best && oo < best_off)) {
            best = ob;
            best_off = oo;
            val = ov;
          }
        }
        if (l == 0) bond_sign[j] = val < 0 ? -1.0 : 1.0;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < g.E * cols; idx += blockDim.x) cur[idx] *= bond_sign[g.fd.div(g.fE.div(idx))];
    __syncthreads();
    double nrm2 = 0;
    for (int i = tid; i < g.E * cols; i += blockDim.x) nrm2 += cur[i] * cur[i];
    nrm2 = block_sum_d(nrm2, red);
    const double sc = (prm.normalize && nrm2 > 0) ? 1.0 / sqrt(nrm2) : 1.0;
    const int ncol = g.d * g.dims[g.j];  // columns of the bond's layout
    for (int idx = tid; idx < g.E * ncol; idx += blockDim.x) {
      const int c = (int)g.fE.div(idx), e = idx - c * g.E;
      const int b = (int)g.fd.div(c), s = c - b * g.d;
      const double val = b < keep ? cur[e + (long long)g.E * (s + g.d * b)] * sc : 0.0;
      psi[g.psi_off + nat_off(g, s, e, b)] = val;
    }
    __syncthreads();
    phase_mark(prm, 6 + side);
  }
}

// One-site gates (apply.jl:110-116): psi_v[s', b] = sum_s O[s', s] psi_v[s, b]
// in place, then psi_v /= ||psi_v|| if normalize (:113-115).  One workgroup per
// gate; the vertices of one launch are distinct.  ops: d x d column-major per
// gate at op_off[i].
__global__ __launch_bounds__(256) void k_apply_site(const int* __restrict__ verts, const VDesc* __restrict__ vds,
                                                    double* __restrict__ psi, const double* __restrict__ ops,
                                                    const long long* __restrict__ op_off, int normalize,
                                                    double* __restrict__ norms) {
  __shared__ double O[AP_SITE_D * AP_SITE_D];
  __shared__ double red[4];
  const VDesc V = vds[verts[blockIdx.x]];
  const int d = V.dims[0];
  const int cols = V.D / d;
  double* t = psi + V.psi_off;
  for (int i = threadIdx.x; i < d * d; i += blockDim.x) O[i] = ops[op_off[blockIdx.x] + i];
  __syncthreads();
  double nrm2 = 0;
  for (int c = threadIdx.x; c < cols; c += blockDim.x) {
    double x[AP_SITE_D], y[AP_SITE_D];
    for (int s = 0; s < d; ++s) x[s] = t[(long long)c * d + s];
    for (int sp = 0; sp < d; ++sp) {
      double acc = 0;
      for (int s = 0; s < d; ++s) acc += O[sp + d * s] * x[s];
      y[sp] = acc;
      nrm2 += acc * acc;
    }
    for (int s = 0; s < d; ++s) t[(long long)c * d + s] = y[s];
  }
  const double tot = block_sum_d(nrm2, red);
  if (threadIdx.x == 0) norms[blockIdx.x] = sqrt(tot);
  if (!normalize || !(tot > 0)) return;
  const double sc = 1.0 / sqrt(tot);
  for (long long i = threadIdx.x; i < V.D; i += blockDim.x) t[i] *= sc;
}

// dynamic LDS of k_apply_gates<false> / <true>: the matrices (LDS variant
// only), then rot, four reflector-head vectors, the singular values, the
// block reduction and (global-memory variant) one environment work area per wave
size_t apply_lds_bytes() {
  return sizeof(double) * (2 * AP_XY + 2 * 1024 + 2 * AP_MAXEXT * 2 * AP_ENV * AP_ENV + 4 * 32 + 4 * AP_N_BIG +
                           AP_TH + 16);
}
size_t apply_big_lds_bytes() {
  return sizeof(double) * (AP_BIG_SMALL + AP_ENV_BIG * AP_ENV_BIG + AP_TH_BIG * AP_TH_BIG);
}

// per-side sizes of a bond for the kernel limits: E (external size), n = d chi_b
// (the bond as it is), cols = d chi (the bond's layout during the call)
struct SideSize {
  long long E, n, cols;
};
SideSize side_size(const itnbp_net* net, int x, int e, const std::vector<int>& chi) {
  SideSize z{1, (long long)net->phys[x] * net->chi[e], (long long)net->phys[x] * chi[e]};
  for (int a = net->inc_ptr[x]; a < net->inc_ptr[x + 1]; ++a)
    if (net->inc_edge[a] != e) z.E *= chi[net->inc_edge[a]];
  return z;
}

// global workspace of k_apply_gates<true> per gate, in doubles, for side
// matrices of q = E n and x = E max(n, cols) doubles
long long big_stride(long long q, long long x) {
  return 2 * q + 2 * x + 2LL * AP_F_BIG + (long long)AP_TH_BIG * AP_TH_BIG +
         2LL * AP_MAXEXT * 2 * AP_ENV_BIG * AP_ENV_BIG + 2LL * AP_N_BIG * AP_N_BIG;
}

// Checks the call and fills the launch items; goff receives the number of gate
// matrix elements.
// Rank bound of Theta for the bond e (apply.jl:74-81): each side's QR keeps
// r = min(E, d chi_e) rows, and Theta is (r_u d_u) x (r_v d_v).
long long theta_rank_bound(const itnbp_net* net, int e) {
  long long rank_bound = 1;
  for (int side = 0; side < 2; ++side) {
    const int x = net->edges[2 * e + side];
    long long E = 1;
    for (int a = net->inc_ptr[x]; a < net->inc_ptr[x + 1]; ++a)
      if (net->inc_edge[a] != e) E *= net->chi[net->inc_edge[a]];
    const long long r = std::min(E, (long long)net->phys[x] * net->chi[e]) * net->phys[x];
    rank_bound = side == 0 ? r : std::min(rank_bound, r);
  }
  return rank_bound;
}

// Checks the call against the kernel's limits with bond dimensions chi (the
// network's, or those it will have after growing) and fills the launch items;
// goff receives the number of gate matrix elements.  With fixed_bonds the
// kept rank must fit the bond as it is.
//
// With big non-null, a bond past the LDS kernel's limits that fits those of
// the global-memory variant is marked (*big)[i] = 1 instead of refused.
int prepare_apply(itnbp_net* net, int n, const int* edge_ids, int maxdim, double cutoff, double* svals,
                  int max_keep, const std::vector<int>& chi, bool fixed_bonds, std::vector<ApplyItem>& items,
                  long long& goff, std::vector<char>* big = nullptr) {
  itnbp_ctx* ctx = net->ctx;
  if (net->dtype != ITNBP_F64) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "apply needs FP64 site tensors");
  if (svals && max_keep <= 0) return itnbp_fail(ctx, ITNBP_ERR_ARG, "max_keep must be > 0 with svals");
  items.assign(n, ApplyItem{});
  if (big) big->assign(n, 0);
  std::vector<char> used(net->nv, 0);
  goff = 0;
  for (int i = 0; i < n; ++i) {
    const int e = edge_ids[i];
    if (e < 0 || e >= net->ne) return itnbp_fail(ctx, ITNBP_ERR_ARG, "edge id out of range");
    const int u = net->edges[2 * e], v = net->edges[2 * e + 1];
    if (used[u] || used[v]) return itnbp_fail(ctx, ITNBP_ERR_ARG, "the bonds of one apply call must be disjoint");
    used[u] = used[v] = 1;
    // without a cutoff the kept rank is min(rank bound of Theta, maxdim), known
    // before the launch: a bond that cannot change must hold it
    const long long rank_bound = theta_rank_bound(net, e);
    const long long keep_bound = maxdim > 0 ? std::min<long long>(rank_bound, maxdim) : rank_bound;
    if (fixed_bonds && cutoff < 0 && keep_bound > chi[e])
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                        "apply: the truncation would keep more values than the bond dimension; pass maxdim <= chi");
    // product environment: every message into u but from v (and into v but from u) must be set
    const char* lds_why = nullptr;  // the first LDS limit the bond exceeds
    for (int side = 0; side < 2; ++side) {
      const int x = side == 0 ? u : v;
      const VDesc& V = net->vdesc[x];
      if (V.k > AP_MAXEXT + 1) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "apply supports degree <= 4");
      int chi_ext = 1;
      for (int a = net->inc_ptr[x]; a < net->inc_ptr[x + 1]; ++a) {
        const int f = net->inc_edge[a];
        if (f == e) continue;
        if (!net->msg_set[msg_id_into(net, f, x)])
          return itnbp_fail(ctx, ITNBP_ERR_STATE, "environment message not set");
        chi_ext = std::max(chi_ext, chi[f]);
      }
      const SideSize z = side_size(net, x, e, chi);
      const long long d = net->phys[x], r = std::min(z.E, z.n);
      if (lds_why) continue;
      if (chi_ext > AP_ENV) lds_why = "apply supports external bond dimensions <= 8";
      else if (z.E * z.n > AP_XY || z.E * z.cols > AP_XY) lds_why = "apply: site tensor too large for the LDS kernel";
      else if (r * d > AP_TH) lds_why = "apply: Theta larger than 32x32";
      else if (r * z.n > 1024) lds_why = "apply: R factor larger than 1024";
    }
    if (lds_why) {
      // the global-memory variant: external bonds <= 32, d chi_b <= 64, Theta <= 128 x 128
      bool fits = big != nullptr;
      for (int side = 0; side < 2 && fits; ++side) {
        const int x = side == 0 ? u : v;
        const SideSize z = side_size(net, x, e, chi);
        for (int a = net->inc_ptr[x]; a < net->inc_ptr[x + 1]; ++a)
          if (net->inc_edge[a] != e && chi[net->inc_edge[a]] > AP_ENV_BIG) fits = false;
        fits = fits && z.n <= AP_N_BIG && std::min(z.E, z.n) * net->phys[x] <= AP_TH_BIG &&
               chi[e] <= AP_KEEP_BIG && z.E * std::max(z.n, z.cols) < (1LL << 30);
      }
      if (!fits) {
        if (!big) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, lds_why);
        return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                          "apply supports external bond dimensions <= 32, d chi <= 64 on the bond and Theta <= "
                          "128 x 128");
      }
      (*big)[i] = 1;
    }
    items[i].u = u;
    items[i].v = v;
    items[i].ju = axis_of(net, u, e);
    items[i].jv = axis_of(net, v, e);
    items[i].gate = goff;
    items[i].slot = i;
    items[i].chib_in = net->chi[e];
    goff += (long long)net->phys[u] * net->phys[v] * net->phys[u] * net->phys[v];
  }
  return ITNBP_OK;
}

// Device memory of one apply call: launch items, gate matrices, the per-gate
// workspace of k_apply_gates (2 AP_SIDE doubles per gate) and the per-gate outputs.
struct ApplyBufs {
  ApplyItem* items;
  double* gates;
  double* work;
  double* te;
  double* sv;
  int* kept;
  int* st;
};

// bytes of each part, in the order of ApplyBufs, every part 256-byte aligned
void apply_buf_sizes(int n, long long goff, int mk, size_t (&sz)[7]) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  sz[0] = al(n * sizeof(ApplyItem));
  sz[1] = al((size_t)goff * 8);
  sz[2] = al((size_t)n * 2 * AP_SIDE * 8);
  sz[3] = al((size_t)n * 8);
  sz[4] = al((size_t)n * mk * 8);
  sz[5] = al((size_t)n * 4);
  sz[6] = al((size_t)n * 4);
}

ApplyBufs carve_apply_bufs(char* base, int n, long long goff, int mk) {
  size_t sz[7];
  apply_buf_sizes(n, goff, mk, sz);
  ApplyBufs b;
  b.items = (ApplyItem*)base;
  base += sz[0];
  b.gates = (double*)base;
  base += sz[1];
  b.work = (double*)base;
  base += sz[2];
  b.te = (double*)base;
  base += sz[3];
  b.sv = (double*)base;
  base += sz[4];
  b.kept = (int*)base;
  base += sz[5];
  b.st = (int*)base;
  return b;
}

// upload, one launch, read-back; every device buffer comes from b
// upload, the launches, read-back; every device buffer but the global-memory
// variant's workspace comes from b.  Gates marked in big run on
// k_apply_gates<true> after the others, in chunks whose workspace fits the
// device's free memory (at most 16 GiB).
int run_apply(itnbp_net* net, int n, const std::vector<ApplyItem>& items_in, const double* gates, long long goff,
              int maxdim, double cutoff, int normalize, double* truncerr, double* svals, int mk, int* kept,
              const ApplyBufs& b, const std::vector<char>* big = nullptr) {
  itnbp_ctx* ctx = net->ctx;
  // the LDS kernel's gates first, then the global-memory variant's (every
  // output is indexed by the gate's slot, so the order is free)
  std::vector<ApplyItem> items;
  items.reserve(n);
  for (int i = 0; i < n; ++i)
    if (!big || !(*big)[i]) items.push_back(items_in[i]);
  const int nl = (int)items.size();
  for (int i = 0; i < n; ++i)
    if (big && (*big)[i]) items.push_back(items_in[i]);
  const int nb = n - nl;
  // one upload of the launch items and the gates, laid out as in b (the gates
  // follow the items at their aligned offset)
  size_t sz[7];
  apply_buf_sizes(n, goff, mk, sz);
  std::vector<char> up(sz[0] + (size_t)goff * 8);
  std::memcpy(up.data(), items.data(), n * sizeof(ApplyItem));
  std::memcpy(up.data() + sz[0], gates, (size_t)goff * 8);
  if ((char*)b.gates != (char*)b.items + sz[0])
    return itnbp_fail(ctx, ITNBP_ERR_HIP, "apply: workspace layout");
  if (hipMemcpyAsync(b.items, up.data(), up.size(), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    return itnbp_fail(ctx, ITNBP_ERR_HIP, "apply: upload failed");
  ApplyParams prm;
  prm.maxdim = maxdim;
  prm.cutoff = cutoff;
  prm.normalize = normalize;
  prm.eig_cutoff = 10 * 2.220446049250313e-16;
  prm.max_keep = mk;
  const size_t lds = apply_lds_bytes();
  hipError_t ea = hipFuncSetAttribute((const void*)k_apply_gates<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds);
  if (ea != hipSuccess)
    return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("apply: LDS attribute: ") + hipGetErrorString(ea));
  DevBuf tsb;
  std::unique_ptr<DevBuf, void (*)(DevBuf*)> ts_guard(&tsb, [](DevBuf* x) { free_buf(*x); });
  if (getenv("ITNBP_APPLY_PHASES")) {
    ITNBP_RETCHK(ensure_buf(ctx, tsb, (size_t)n * 16 * sizeof(unsigned long long)));
    ITNBP_HIPCHK(ctx, hipMemsetAsync(tsb.p, 0, (size_t)n * 16 * sizeof(unsigned long long), ctx->stream));
    prm.ts = (unsigned long long*)tsb.p;
  }
  DevBuf bw;
  std::unique_ptr<DevBuf, void (*)(DevBuf*)> bw_guard(&bw, [](DevBuf* x) { free_buf(*x); });
  // the global-memory variant's workspace first, so that a failed allocation
  // leaves every tensor of the call untouched
  ApplyParams pb = prm;
  int chunk = 1;
  size_t lds_big = 0;
  if (nb > 0) {
    // side matrices of the largest gate: E n for the QR, E max(n, d chi) for the work
    long long q = 1, xs = 1;
    for (int i = nl; i < n; ++i)
      for (int side = 0; side < 2; ++side) {
        const VDesc& V = net->vdesc[side == 0 ? items[i].u : items[i].v];
        const int j = side == 0 ? items[i].ju : items[i].jv;
        long long E = 1;
        for (int a = 1; a <= V.k; ++a)
          if (a != j) E *= V.dims[a];
        const long long nn = (long long)V.dims[0] * items[i].chib_in, cols = (long long)V.dims[0] * V.dims[j];
        q = std::max(q, E * nn);
        xs = std::max(xs, E * std::max(nn, cols));
      }
    pb.big_q = q;
    pb.big_x = xs;
    pb.big_stride = (big_stride(q, xs) + 31) & ~31LL;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    const size_t budget = std::min<size_t>(fr / 2, 16ULL << 30);
    const long long per = pb.big_stride * 8;
    chunk = (int)std::max<long long>(1, std::min<long long>(nb, (long long)(budget / (size_t)per)));
    ITNBP_RETCHK(ensure_buf(ctx, bw, (size_t)chunk * per));
    lds_big = apply_big_lds_bytes();
    hipError_t eb = hipFuncSetAttribute((const void*)k_apply_gates<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds_big);
    if (eb != hipSuccess)
      return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("apply: LDS attribute: ") + hipGetErrorString(eb));
  }
  const int pm0 = prof_mark(ctx);
  if (nl > 0)
    hipLaunchKernelGGL(k_apply_gates<false>, dim3(nl), dim3(256), lds, ctx->stream, (const ApplyItem*)b.items,
                       net->d_vdesc, (double*)net->d_psi, net->d_msg[net->cur], (const double*)b.gates, prm,
                       b.work, b.te, svals ? b.sv : nullptr, b.kept, b.st);
  hipError_t ek = hipGetLastError();
  prof_add(ctx, "apply_gates", pm0, prof_mark(ctx));
  if (ek == hipSuccess && nb > 0) {
    const int pm1 = prof_mark(ctx);
    for (int c0 = nl; c0 < n && ek == hipSuccess; c0 += chunk) {
      const int cn = std::min(chunk, n - c0);
      pb.ts = prm.ts ? prm.ts + (size_t)c0 * 16 : nullptr;
      hipLaunchKernelGGL(k_apply_gates<true>, dim3(cn), dim3(256), lds_big, ctx->stream,
                         (const ApplyItem*)b.items + c0, net->d_vdesc, (double*)net->d_psi, net->d_msg[net->cur],
                         (const double*)b.gates, pb, (double*)bw.p, b.te, svals ? b.sv : nullptr, b.kept, b.st);
      ek = hipGetLastError();
    }
    prof_add(ctx, "apply_gates_global", pm1, prof_mark(ctx));
  }
  if (prm.ts && ek == hipSuccess) {
    // mean duration of every phase over the gates that ran to the end, in us,
    // for the LDS kernel's gates [0, nl) and the global-memory variant's [nl, n)
    std::vector<unsigned long long> ts((size_t)n * 16);
    int khz = 0;
    ek = hipMemcpyAsync(ts.data(), prm.ts, ts.size() * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (ek == hipSuccess) ek = hipStreamSynchronize(ctx->stream);
    if (ek == hipSuccess) ek = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device);
    for (int part = 0; part < 2 && ek == hipSuccess && khz > 0; ++part) {
      const int i0 = part == 0 ? 0 : nl, i1 = part == 0 ? nl : n;
      double mean[7] = {0, 0, 0, 0, 0, 0, 0}, sweeps = 0, tsq = 0;
      int cnt = 0;
      for (int i = i0; i < i1; ++i) {
        const unsigned long long* t = &ts[(size_t)i * 16];
        if (!t[7]) continue;
        ++cnt;
        for (int p = 0; p < 7; ++p) mean[p] += (double)(t[p + 1] - t[p]) * 1e3 / khz;
        sweeps += (double)t[8];
        tsq += (double)t[9];
      }
      if (cnt == 0) continue;
      const double c = cnt;
      double sub[3] = {0, 0, 0}, load = 0;
      for (int i = i0; i < i1; ++i) {
        const unsigned long long* t = &ts[(size_t)i * 16];
        if (!t[7] || !t[13]) continue;
        load += (double)(t[10] - t[1]) * 1e3 / khz;
        for (int p = 0; p < 3; ++p) sub[p] += (double)(t[11 + p] - t[10 + p]) * 1e3 / khz;
      }
      const char* what = part == 0 ? "LDS kernel" : "global-memory variant";
      fprintf(stderr, "  %s, side u of qr (us): psi load %.1f, mode products %.1f, factorisation %.1f, save %.1f\n",
              what, load / c, sub[0] / c, sub[1] / c, sub[2] / c);
      fprintf(stderr,
              "apply phases, %s (us, mean over %d gates): roots %.1f, qr %.1f (%.2f tall-skinny sides), theta %.1f, "
              "svd %.1f (%.1f sweeps), factors %.1f, rebuild_u %.1f, rebuild_v %.1f\n",
              what, cnt, mean[0] / c, mean[1] / c, tsq / c, mean[2] / c, mean[3] / c, sweeps / c, mean[4] / c,
              mean[5] / c, mean[6] / c);
    }
  }
  std::vector<int> st(n);
  std::vector<int> kp(n);
  // one read-back of the outputs (truncation errors, singular values, kept
  // ranks, status: consecutive in b)
  std::vector<char> dn(sz[3] + sz[4] + sz[5] + (size_t)n * 4);
  if (ek == hipSuccess && ((char*)b.sv != (char*)b.te + sz[3] || (char*)b.kept != (char*)b.sv + sz[4] ||
                           (char*)b.st != (char*)b.kept + sz[5]))
    return itnbp_fail(ctx, ITNBP_ERR_HIP, "apply: workspace layout");
  if (ek == hipSuccess) ek = hipMemcpyAsync(dn.data(), b.te, dn.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (ek == hipSuccess) ek = hipStreamSynchronize(ctx->stream);
  if (ek != hipSuccess) return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("apply: ") + hipGetErrorString(ek));
  std::memcpy(truncerr, dn.data(), (size_t)n * 8);
  if (svals) std::memcpy(svals, dn.data() + sz[3], (size_t)n * mk * 8);
  std::memcpy(kp.data(), dn.data() + sz[3] + sz[4], (size_t)n * 4);
  std::memcpy(st.data(), dn.data() + sz[3] + sz[4] + sz[5], (size_t)n * 4);
  // a gate the kernel refused (its bond untouched) reports kept = -1
  for (int i = 0; i < n; ++i)
    if (st[i]) kp[i] = -1;
  if (kept) std::copy(kp.begin(), kp.end(), kept);
  invalidate_fast_state(net);
  for (int i = 0; i < n; ++i)
    if (st[i] == 2)
      return itnbp_fail(ctx, ITNBP_ERR_NONFINITE,
                        "apply: non-finite two-site tensor (that bond's tensors are unchanged; the other gates of "
                        "the call were applied)");
  for (int i = 0; i < n; ++i)
    if (st[i])
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                        "apply: the truncation keeps more values than the bond dimension (that bond's tensors "
                        "are unchanged; the other gates of the call were applied); pass maxdim <= chi");
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(truncerr[i])) return itnbp_fail(ctx, ITNBP_ERR_NONFINITE, "apply: non-finite result");
  return ITNBP_OK;
}

// ---- workspace of itnbp_apply_gates_noalloc --------------------------------
// At most NOALLOC_MAX_GATES gates per call, gate matrices of at most
// NOALLOC_MAX_GATE_ELEMS elements ((d_u d_v)^2 with d <= 4) and at most
// NOALLOC_MAX_KEEP singular values per gate: 18.4 MiB in all, held by every
// GPU the library runs on until the process ends.
constexpr int NOALLOC_MAX_GATES = 128;
constexpr int NOALLOC_MAX_GATE_ELEMS = 256;
constexpr int NOALLOC_MAX_KEEP = 32;

constexpr size_t noalloc_bytes() {
  return ((NOALLOC_MAX_GATES * sizeof(ApplyItem) + 255) & ~(size_t)255) +
         (size_t)NOALLOC_MAX_GATES * NOALLOC_MAX_GATE_ELEMS * 8 + (size_t)NOALLOC_MAX_GATES * 2 * AP_SIDE * 8 +
         (size_t)NOALLOC_MAX_GATES * 8 + (size_t)NOALLOC_MAX_GATES * NOALLOC_MAX_KEEP * 8 +
         2 * (size_t)NOALLOC_MAX_GATES * 4 + 8 * 256;
}

// The array's name is a host-side stand-in: host code uses only the device
// address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[noalloc_bytes()];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
int g_noalloc_rc[NOALLOC_MAX_DEVICES] = {};
bool g_noalloc_done[NOALLOC_MAX_DEVICES] = {};

}  // namespace

// Resolves the workspace address on `device` (current) once per device; called
// by itnbp_create.  Returns ITNBP_OK or ITNBP_ERR_HIP (then the no-allocation
// entry point refuses every call on that device).
int apply_noalloc_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return ITNBP_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_done[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      ptr = nullptr;
    }
    g_noalloc_ptr[device] = ptr;
    g_noalloc_rc[device] = e == hipSuccess ? ITNBP_OK : ITNBP_ERR_HIP;
    g_noalloc_done[device] = true;
  }
  return g_noalloc_rc[device];
}

extern "C" int itnbp_apply_gates(itnbp_net* net, int n, const int* edge_ids, const double* gates,
                                 int maxdim, double cutoff, int normalize, double* truncerr,
                                 double* svals, int max_keep, int* kept) {
  if (!net || n < 0 || (n > 0 && (!edge_ids || !gates || !truncerr))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  itnbp_ctx* ctx = net->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    // The new bond has the kept dimension (apply.jl:81-93).  A bond whose kept
    // rank may exceed its dimension first grows (zero-padded) to the largest
    // rank the truncation can keep, min(rank bound of Theta, maxdim); the kernel
    // writes the kept columns followed by zeros; then every bond of the call is
    // cut to its kept rank, which drops only zeros.
    for (int i = 0; i < n; ++i)
      if (edge_ids[i] < 0 || edge_ids[i] >= net->ne) return itnbp_fail(ctx, ITNBP_ERR_ARG, "edge id out of range");
    const std::vector<int> chi0 = net->chi;
    std::vector<int> target(n), chi_grow = net->chi;
    for (int i = 0; i < n; ++i) {
      const int e = edge_ids[i];
      const long long rb = theta_rank_bound(net, e);
      target[i] = (int)std::min<long long>(maxdim > 0 ? std::min<long long>(rb, maxdim) : rb, 64);
      chi_grow[e] = std::max(chi_grow[e], target[i]);
    }
    std::vector<ApplyItem> items;
    std::vector<char> big;
    long long goff = 0;
    ITNBP_RETCHK(
        prepare_apply(net, n, edge_ids, maxdim, cutoff, svals, max_keep, chi_grow, false, items, goff, &big));
    ITNBP_RETCHK(net_resize_bonds(net, chi_grow));
    const int mk = svals ? max_keep : 1;
    size_t sz[7];
    apply_buf_sizes(n, goff, mk, sz);
    size_t total = 0;
    for (size_t b : sz) total += b;
    // the net keeps the workspace between calls (a gate round is several calls)
    ITNBP_RETCHK(ensure_buf(ctx, net->apply_ws, total));
    std::vector<int> kp(n, -1);
    const int rc = run_apply(net, n, items, gates, goff, maxdim, cutoff, normalize, truncerr, svals, mk, kp.data(),
                             carve_apply_bufs((char*)net->apply_ws.p, n, goff, mk), &big);
    std::vector<int> chi_cut = net->chi;
    if (rc != ITNBP_OK) {
      // the gates that ran keep their new bonds; a refused gate's bond (its
      // tensors untouched, grown with zeros) returns to its old dimension
      for (int i = 0; i < n; ++i)
        chi_cut[edge_ids[i]] = kp[i] < 0 ? chi0[edge_ids[i]] : std::max(1, std::min(kp[i], target[i]));
      if (kept) std::copy(kp.begin(), kp.end(), kept);
      (void)net_resize_bonds(net, chi_cut);
      return rc;
    }
    for (int i = 0; i < n; ++i) {
      kp[i] = std::max(1, std::min(kp[i], target[i]));
      chi_cut[edge_ids[i]] = kp[i];
      if (svals)
        for (int j = kp[i]; j < mk; ++j) svals[(size_t)i * mk + j] = 0.0;
    }
    if (kept) std::copy(kp.begin(), kp.end(), kept);
    return net_resize_bonds(net, chi_cut);
  });
}

extern "C" int itnbp_apply_site_gates(itnbp_net* net, int n, const int* verts, const double* ops, int normalize,
                                      double* norms) {
  if (!net || n < 0 || (n > 0 && (!verts || !ops))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  itnbp_ctx* ctx = net->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    if (net->dtype != ITNBP_F64) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "apply needs FP64 site tensors");
    std::vector<char> used(net->nv, 0);
    std::vector<long long> off(n);
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
      const int v = verts[i];
      if (v < 0 || v >= net->nv) return itnbp_fail(ctx, ITNBP_ERR_ARG, "vertex out of range");
      if (used[v]) return itnbp_fail(ctx, ITNBP_ERR_ARG, "the vertices of one site-gate call must be distinct");
      used[v] = 1;
      const int d = net->phys[v];
      if (d > AP_SITE_D) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "one-site gates support d <= 16");
      off[i] = tot;
      tot += (long long)d * d;
    }
    DevBuf b;
    const size_t ob = ((size_t)n * 4 + 255) & ~(size_t)255, oo = ((size_t)n * 8 + 255) & ~(size_t)255;
    ITNBP_RETCHK(ensure_buf(ctx, b, ob + 2 * oo + (size_t)tot * 8));
    char* p = (char*)b.p;
    int* d_v = (int*)p;
    long long* d_off = (long long*)(p + ob);
    double* d_nrm = (double*)(p + ob + oo);
    double* d_ops = (double*)(p + ob + 2 * oo);
    std::vector<double> nrm(n);
    hipError_t e = hipMemcpyAsync(d_v, verts, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, off.data(), (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_ops, ops, (size_t)tot * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
      const int pm0 = prof_mark(ctx);
      hipLaunchKernelGGL(k_apply_site, dim3(n), dim3(256), 0, ctx->stream, (const int*)d_v, net->d_vdesc,
                         (double*)net->d_psi, (const double*)d_ops, (const long long*)d_off, normalize, d_nrm);
      e = hipGetLastError();
      prof_add(ctx, "apply_site", pm0, prof_mark(ctx));
    }
    if (e == hipSuccess) e = hipMemcpyAsync(nrm.data(), d_nrm, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    free_buf(b);
    if (e != hipSuccess) return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("apply_site_gates: ") + hipGetErrorString(e));
    invalidate_fast_state(net);
    if (norms) std::copy(nrm.begin(), nrm.end(), norms);
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(nrm[i])) return itnbp_fail(ctx, ITNBP_ERR_NONFINITE, "apply_site_gates: non-finite result");
    return ITNBP_OK;
  });
}

extern "C" int itnbp_apply_gates_noalloc(itnbp_net* net, int n, const int* edge_ids, const double* gates,
                                         int maxdim, double cutoff, int normalize, double* truncerr,
                                         double* svals, int max_keep, int* kept) {
  if (!net || n < 0 || (n > 0 && (!edge_ids || !gates || !truncerr))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  itnbp_ctx* ctx = net->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    if (n > NOALLOC_MAX_GATES)
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                        "apply_gates_noalloc: at most " + std::to_string(NOALLOC_MAX_GATES) + " gates per call");
    const int mk = svals ? max_keep : 1;
    if (mk > NOALLOC_MAX_KEEP)
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                        "apply_gates_noalloc: max_keep at most " + std::to_string(NOALLOC_MAX_KEEP));
    std::vector<ApplyItem> items;
    long long goff = 0;
    // no allocation, so no resizing: every bond keeps its dimension and fewer
    // kept values are zero-padded
    ITNBP_RETCHK(prepare_apply(net, n, edge_ids, maxdim, cutoff, svals, max_keep, net->chi, true, items, goff));
    if (goff > (long long)n * NOALLOC_MAX_GATE_ELEMS)
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                        "apply_gates_noalloc: gate matrices larger than (d_u d_v)^2 = 256 elements");
    void* base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_noalloc_mu);
      if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES && g_noalloc_done[ctx->device])
        base = g_noalloc_ptr[ctx->device];
    }
    if (!base)
      return itnbp_fail(ctx, ITNBP_ERR_HIP, "apply_gates_noalloc: the workspace address of this device is unknown");
    size_t sz[7];
    apply_buf_sizes(n, goff, mk, sz);
    size_t total = 0;
    for (size_t b : sz) total += b;
    if (total > noalloc_bytes())
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "apply_gates_noalloc: call exceeds the workspace");
    return run_apply(net, n, items, gates, goff, maxdim, cutoff, normalize, truncerr, svals, mk, kept,
                     carve_apply_bufs((char*)base, n, goff, mk));
  });
}

// ============================================================================
// Adjoint of the gate application (vector-Jacobian product of
// simple_update_bp, src/apply.jl:33-95, with the normalisation of :134-137).
//
// The kernel recomputes each bond's forward pass in a gauge of its own and
// pulls the cotangents of psi_u', psi_v' back to psi_u, psi_v, the gate and
// every environment message.  With X = psi (x) sqrt(env) as the E x n matrix of
// a side (E external, n = d chi_b) the QR factor only enters through a gauge
// that cancels between Theta and the rebuild, so:
//   - a tall side (E >= n) uses R = C^{1/2}, C = X^T X, and the rebuild factor
//     W = psi_P C^{-1/2}, psi_P = psi (x) P with P the projector on the kept
//     environment eigenvectors (S Si = P);
//   - a wide side (E < n) uses R = X and W = (x) Si.
// Theta = G (R_u R_v), thin SVD U S V^T, F_u = U_k sqrt(S_k), F_v = V_k sqrt(S_k),
// psi_u' = W_u F_u times the bond signs of the forward gauge, then normalised.
// Pullbacks, in reverse: normalisation; W F products; sqrt(S); the thin SVD
// (full-rank formula with the kept/discarded couplings, 1/(s_j^2 - s_i^2)
// broadened by 1e-12 s_max^2 against degenerate values); the contraction with
// the gate; R = C^{1/2} and C^{-1/2} through the eigenbasis of C with the
// divided differences 1/(s_i + s_j) and -1/(s_i s_j (s_i + s_j)); X = psi (x) S;
// sqrt and inverse sqrt of every environment through its eigenbasis (kept
// eigenvalues; a dropped one contributes nothing).  P is taken as constant.
// One workgroup per bond, FP64, intermediates in a global workspace per bond.
// ============================================================================
namespace {

constexpr int VJ_SM = 1024;
constexpr int VJ_ENV = 64 * 6 + 16;  // per external axis: V, S, Si, Sb, Sib, tmp (8x8 each) + w
// per side: 9 matrices of up to AP_XY, 8 of up to 1024, 3 environments
constexpr long long VJ_SIDE = 9LL * AP_XY + 8LL * VJ_SM + AP_MAXEXT * VJ_ENV + 64;
// shared: 14 matrices of up to 1024 and small vectors
constexpr long long VJ_SHARED = 14LL * VJ_SM + 6 * 64;
constexpr long long VJ_GATE = 2 * VJ_SIDE + VJ_SHARED;

struct VjpIo {
  long long out_u, out_v;  // cotangents of psi_u', psi_v' (new bond dimension keep)
  long long in_u, in_v;    // cotangents of psi_u, psi_v
  long long dgate;         // cotangent of the gate
  int keep, pad;
};

__device__ inline void bsync() {
  __threadfence_block();
  __syncthreads();
}

// C (m x n, ld ldc) = op(A) op(B) (+ C if acc), column-major; A m x k (or k x m
// transposed), B k x n (or n x k transposed)
__device__ void gmm(int m, int n, int k, const double* A, int lda, bool ta, const double* B, int ldb, bool tb,
                    double* C, int ldc, bool acc = false) {
  for (int idx = threadIdx.x; idx < m * n; idx += blockDim.x) {
    const int i = idx % m, j = idx / m;
    double s = 0;
    for (int l = 0; l < k; ++l) {
      const double a = ta ? A[l + (long long)lda * i] : A[i + (long long)lda * l];
      const double b = tb ? B[j + (long long)ldb * l] : B[l + (long long)ldb * j];
      s += a * b;
    }
    C[i + (long long)ldc * j] = acc ? C[i + (long long)ldc * j] + s : s;
  }
  bsync();
}

// Y = X with the symmetric chi x chi matrix Mt (row-major) on external axis pa
__device__ void ext_mp_g(const SideGeom& g, int pa, const double* X, double* Y, int cols, const double* Mt) {
  const int stride = (int)g.fst[pa].d;
  const int chi = g.dims[g.ext[pa]];
  for (int idx = threadIdx.x; idx < g.E * cols; idx += blockDim.x) {
    const int c = (int)g.fE.div(idx), e = idx - c * g.E;
    const int q = (int)g.fst[pa].div(e), lo = e - q * stride;
    const int hi = (int)g.fchi[pa].div(q), b2 = q - hi * chi;
    double acc = 0;
    for (int b = 0; b < chi; ++b) acc += X[(lo + stride * (b + chi * hi)) + (long long)g.E * c] * Mt[b * chi + b2];
    Y[idx] = acc;
  }
  bsync();
}

// Symmetric eigendecomposition of one chi x chi message by one wave (the Jacobi
// of env_roots_wave): V (row-major, V[r * chi + m] = component r of eigenvector
// m), w[m], and the kept mask of map_eigvals' truncation (cutoff on |w|).
__device__ unsigned env_eig_wave(const double* M, int chi, double cutoff, double* Vout, double* wout) {
  const int l = threadIdx.x & 63;
  const int ne = chi + (chi & 1);
  const bool live = l < ne * ne;
  const int i = live ? l / ne : 0, j = live ? l % ne : 0;
  const bool in = live && i < chi && j < chi;
  double a = in ? 0.5 * (M[i + chi * j] + M[j + chi * i]) : 0.0;
  double v = (live && i == j) ? 1.0 : 0.0;
  if (chi > 1) {
    auto rot = [&](int p, int q, double& c, double& sn) {
      const double app = __shfl(a, p * ne + p), aqq = __shfl(a, q * ne + q), apq = __shfl(a, p * ne + q);
      c = 1.0;
      sn = 0.0;
      if (q < chi && apq != 0.0) {
        const double zeta = (aqq - app) / (2 * apq);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
        c = 1 / sqrt(1 + t * t);
        sn = c * t;
      }
    };
    double prev = INFINITY;
    for (int sweep = 0; sweep < 40; ++sweep) {
      double off = (in && i != j) ? a * a : 0.0, diag = (in && i == j) ? a * a : 0.0;
      for (int m = 32; m > 0; m >>= 1) {
        off += __shfl_xor(off, m);
        diag += __shfl_xor(diag, m);
      }
      // converged, or the off-diagonal mass no longer decreases (rounding level)
      if (off <= 1e-30 * diag || off == 0.0 || off >= prev) break;
      prev = off;
      for (int st = 0; st < ne - 1; ++st) {
        // (x - 1 - st) lies in (-(ne - 1), ne - 1) and (pp - 1 + st) in [0, 2 (ne - 1)):
        // conditional corrections instead of modulos
        auto partner = [&](int x) {
          int t = x - 1 - st;
          if (t < 0) t += ne - 1;
          const int pos = x == 0 ? 0 : 1 + t;
          const int pp = ne - 1 - pos;
          int u = pp - 1 + st;
          if (u >= ne - 1) u -= ne - 1;
          return pp == 0 ? 0 : 1 + u;
        };
        const int ri = partner(i), cj = partner(j);
        double cr, sr, cc, sc;
        rot(i < ri ? i : ri, i < ri ? ri : i, cr, sr);
        rot(j < cj ? j : cj, j < cj ? cj : j, cc, sc);
        const double ao = __shfl(a, ri * ne + j);
        const double ar = i < ri ? cr * a - sr * ao : sr * ao + cr * a;
        const double aro = __shfl(ar, i * ne + cj), vo = __shfl(v, i * ne + cj);
        a = j < cj ? cc * ar - sc * aro : sc * aro + cc * ar;
        v = j < cj ? cc * v - sc * vo : sc * vo + cc * v;
      }
    }
  }
  double w[AP_ENV];
#pragma unroll
  for (int m = 0; m < AP_ENV; ++m) w[m] = __shfl(a, (m < ne ? m : 0) * ne + (m < ne ? m : 0));
  int order[AP_ENV];
  for (int m = 0; m < chi; ++m) order[m] = m;
  for (int x = 0; x < chi; ++x)
    for (int y = x + 1; y < chi; ++y)
      if (fabs(w[order[y]]) > fabs(w[order[x]])) {
        const int t = order[x];
        order[x] = order[y];
        order[y] = t;
      }
  double tot = 0;
  for (int m = 0; m < chi; ++m) tot += fabs(w[m]);
  int keep = chi;
  double disc = 0;
  while (keep > 1 && tot > 0 && (disc + fabs(w[order[keep - 1]])) / tot <= cutoff) {
    disc += fabs(w[order[keep - 1]]);
    --keep;
  }
  unsigned kmask = 0;
  for (int m = 0; m < keep; ++m) kmask |= 1u << order[m];
  if (in) Vout[i * chi + j] = v;
  if (l < chi) wout[l] = w[l];
  return kmask;
}

// S = V f(w) V^T over the kept eigenvalues (row-major chi x chi), f = sqrt
// (which = 0), 1/sqrt (1) or the indicator (2)
__device__ void env_fun(const double* V, const double* w, unsigned kmask, int chi, int which, double* out) {
  for (int idx = threadIdx.x & 63; idx < chi * chi; idx += 64) {
    const int r = idx / chi, c = idx % chi;
    double s = 0;
    for (int m = 0; m < chi; ++m)
      if ((kmask >> m) & 1u) {
        const double f = which == 0 ? sqrt(w[m]) : which == 1 ? 1.0 / sqrt(w[m]) : 1.0;
        s += V[r * chi + m] * V[c * chi + m] * f;
      }
    out[idx] = s;
  }
}

// Mbar (row-major, chi x chi, symmetrised) = V [(V^T Sb V) o K1 + (V^T Sib V) o K2] V^T
// with the divided differences of sqrt and 1/sqrt on the kept eigenvalues
__device__ void env_pullback(const double* V, const double* w, unsigned kmask, int chi, const double* Sb,
                             const double* Sib, double* tmp, double* Mb) {
  const int l = threadIdx.x & 63;
  // tmp = V^T (Sb_sym) V o K1 + V^T (Sib_sym) V o K2   (eigenbasis, row-major)
  for (int idx = l; idx < chi * chi; idx += 64) {
    const int p = idx / chi, q = idx % chi;
    double a = 0, b = 0;
    for (int r = 0; r < chi; ++r)
      for (int c = 0; c < chi; ++c) {
        const double vv = V[r * chi + p] * V[c * chi + q];
        a += vv * 0.5 * (Sb[r * chi + c] + Sb[c * chi + r]);
        b += vv * 0.5 * (Sib[r * chi + c] + Sib[c * chi + r]);
      }
    const bool kp = (kmask >> p) & 1u, kq = (kmask >> q) & 1u;
    double k1 = 0, k2 = 0;
    if (kp && kq) {
      const double sp = sqrt(w[p]), sq = sqrt(w[q]);
      k1 = 1.0 / (sp + sq);
      k2 = -1.0 / (sp * sq * (sp + sq));
    } else if (kp || kq) {
      // one kept (f(x)), one dropped (f = 0): (f(w_p) - f(w_q)) / (w_p - w_q)
      const double wk = kp ? w[p] : w[q], wd = kp ? w[q] : w[p];
      const double den = wk - wd;
      if (den != 0.0) {
        k1 = sqrt(wk) / den;
        k2 = 1.0 / sqrt(wk) / den;
      }
    }
    tmp[idx] = a * k1 + b * k2;
  }
  __threadfence_block();
  for (int idx = l; idx < chi * chi; idx += 64) {
    const int r = idx / chi, c = idx % chi;
    double s = 0, t = 0;
    for (int p = 0; p < chi; ++p)
      for (int q = 0; q < chi; ++q) {
        s += V[r * chi + p] * tmp[p * chi + q] * V[c * chi + q];
        t += V[c * chi + p] * tmp[p * chi + q] * V[r * chi + q];
      }
    Mb[idx] = 0.5 * (s + t);
  }
}

__global__ __launch_bounds__(256) void k_apply_vjp(const ApplyItem* __restrict__ items, const VjpIo* __restrict__ io,
                                                   const VDesc* __restrict__ vds, const double* __restrict__ psi,
                                                   const double* __restrict__ msg, const double* __restrict__ gates,
                                                   ApplyParams prm, const double* __restrict__ dout,
                                                   double* __restrict__ din, double* __restrict__ dgate,
                                                   double* __restrict__ dmsg, double* __restrict__ work,
                                                   int* __restrict__ status) {
  __shared__ double red[8];
  __shared__ int order[AP_TH];
  __shared__ int sh_keep;
  __shared__ unsigned kmask[2][AP_MAXEXT];
  const ApplyItem it = items[blockIdx.x];
  const VjpIo o = io[blockIdx.x];
  const int tid = threadIdx.x;
  SideGeom gs[2] = {side_geom(vds[it.u], it.ju), side_geom(vds[it.v], it.jv)};
  const int vtx[2] = {it.u, it.v};
  double* base = work + (long long)blockIdx.x * VJ_GATE;
  // per side areas
  double *X[2], *Pt[2], *A[2], *W[2], *Xb[2], *Pb[2], *Yb[2], *T1[2], *T2[2];
  double *Z[2], *R[2], *B[2], *Rb[2], *Bb[2], *Cb[2], *Sg[2], *Wb[2], *env[2];
  for (int s = 0; s < 2; ++s) {
    double* p = base + s * VJ_SIDE;
    X[s] = p; Pt[s] = p + AP_XY; A[s] = p + 2 * AP_XY; W[s] = p + 3 * AP_XY; Xb[s] = p + 4 * AP_XY;
    Pb[s] = p + 5 * AP_XY; Yb[s] = p + 6 * AP_XY; T1[s] = p + 7 * AP_XY; T2[s] = p + 8 * AP_XY;
    double* q = p + 9LL * AP_XY;
    Z[s] = q; R[s] = q + VJ_SM; B[s] = q + 2 * VJ_SM; Rb[s] = q + 3 * VJ_SM; Bb[s] = q + 4 * VJ_SM;
    Cb[s] = q + 5 * VJ_SM; Sg[s] = q + 6 * VJ_SM; Wb[s] = q + 7 * VJ_SM;
    env[s] = q + 8 * VJ_SM;
  }
  double* sh = base + 2 * VJ_SIDE;
  double *Th = sh, *Tt = sh + VJ_SM, *As = sh + 2 * VJ_SM, *Vs = sh + 3 * VJ_SM, *Fu = sh + 4 * VJ_SM,
         *Fv = sh + 5 * VJ_SM, *Fub = sh + 6 * VJ_SM, *Fvb = sh + 7 * VJ_SM, *Thb = sh + 8 * VJ_SM,
         *Ttb = sh + 9 * VJ_SM, *Um = sh + 10 * VJ_SM, *Vm = sh + 11 * VJ_SM, *M1 = sh + 12 * VJ_SM,
         *Q2 = sh + 13 * VJ_SM;
  double* svec = sh + 14 * VJ_SM;  // thin singular values (sorted)
  double* sbar = svec + 64;
  double* sgn = sbar + 64;
  double* rot = sgn + 64;  // 128
  auto env_p = [&](int s, int pa, int which) { return env[s] + pa * VJ_ENV + which * 64; };
  // env slots: 0 V, 1 S, 2 Si, 3 Sbar, 4 Sibar, 5 tmp; w at 6 * 64

  // ---- forward -------------------------------------------------------------
  // F1. eigendecompositions of the environments, one wave each
  {
    const int wv = tid >> 6;
    for (int q = wv; q < gs[0].next + gs[1].next; q += (int)blockDim.x / 64) {
      const int s = q < gs[0].next ? 0 : 1;
      const int pa = s == 0 ? q : q - gs[0].next;
      const SideGeom& g = gs[s];
      const int a = g.ext[pa];
      const int chi = g.dims[a];
      double* V = env_p(s, pa, 0);
      double* w = env[s] + pa * VJ_ENV + 6 * 64;
      const unsigned km = env_eig_wave(msg + vds[vtx[s]].in_msg[a], chi, prm.eig_cutoff, V, w);
      __threadfence_block();
      env_fun(V, w, km, chi, 0, env_p(s, pa, 1));
      env_fun(V, w, km, chi, 1, env_p(s, pa, 2));
      env_fun(V, w, km, chi, 2, env_p(s, pa, 5));  // P in the tmp slot until the pullback
      for (int i = (tid & 63); i < 64; i += 64) {
        env_p(s, pa, 3)[i] = 0.0;
        env_p(s, pa, 4)[i] = 0.0;
      }
      if ((tid & 63) == 0) kmask[s][pa] = km;
    }
    bsync();
  }
  // F2-F3. X = psi (x) S, Pt = psi (x) P, R and W per side
  for (int s = 0; s < 2; ++s) {
    const SideGeom& g = gs[s];
    for (int idx = tid; idx < g.E * g.n; idx += blockDim.x) {
      const int e = idx % g.E, c = idx / g.E;
      const double x = psi[g.psi_off + nat_off(g, c % g.d, e, c / g.d)];
      X[s][idx] = x;
      Pt[s][idx] = x;
    }
    bsync();
    for (int pa = 0; pa < g.next; ++pa) {
      ext_mp_g(g, pa, X[s], T1[s], g.n, env_p(s, pa, 1));
      for (int i = tid; i < g.E * g.n; i += blockDim.x) X[s][i] = T1[s][i];
      ext_mp_g(g, pa, Pt[s], T1[s], g.n, env_p(s, pa, 5));
      for (int i = tid; i < g.E * g.n; i += blockDim.x) Pt[s][i] = T1[s][i];
      bsync();
    }
    if (g.E >= g.n) {
      // tall: singular values and right vectors of X give C = Z diag(sigma^2) Z^T
      for (int i = tid; i < g.E * g.n; i += blockDim.x) A[s][i] = X[s][i];
      bsync();
      jacobi_svd(A[s], g.E, g.n, Z[s], rot, red);
      bsync();
      for (int j = tid; j < g.n; j += blockDim.x) {
        double t = 0;
        for (int i = 0; i < g.E; ++i) t += A[s][i + (long long)g.E * j] * A[s][i + (long long)g.E * j];
        Sg[s][j] = sqrt(t);
      }
      bsync();
      for (int idx = tid; idx < g.n * g.n; idx += blockDim.x) {
        const int i = idx % g.n, j = idx / g.n;
        double r = 0, b = 0;
        for (int m = 0; m < g.n; ++m) {
          const double zz = Z[s][i + g.n * m] * Z[s][j + g.n * m];
          r += zz * Sg[s][m];
          b += Sg[s][m] > 0 ? zz / Sg[s][m] : 0.0;
        }
        R[s][idx] = r;  // C^{1/2}, r = n rows
        B[s][idx] = b;  // C^{-1/2}
      }
      bsync();
      gmm(g.E, g.n, g.n, Pt[s], g.E, false, B[s], g.n, false, W[s], g.E);  // W = Pt C^{-1/2}
    } else {
      // wide: R = X (r = E rows), W = (x) Si as an E x E matrix
      for (int i = tid; i < g.E * g.n; i += blockDim.x) R[s][i] = X[s][i];
      for (int idx = tid; idx < g.E * g.E; idx += blockDim.x) {
        const int e = idx % g.E, f = idx / g.E;
        double w = 1.0;
        int re = e, rf = f;
        for (int pa = 0; pa < g.next; ++pa) {
          const int chi = g.dims[g.ext[pa]];
          w *= env_p(s, pa, 2)[(re % chi) * chi + (rf % chi)];
          re /= chi;
          rf /= chi;
        }
        W[s][idx] = w;
      }
      bsync();
    }
  }
  const int du = gs[0].d, dv = gs[1].d;
  const int ru = gs[0].E >= gs[0].n ? gs[0].n : gs[0].E, rv = gs[1].E >= gs[1].n ? gs[1].n : gs[1].E;
  const int nu = gs[0].n, nv = gs[1].n, chib = gs[0].chib;
  const int Mr = ru * du, Nc = rv * dv, dd = du * dv;
  const double* G = gates + it.gate;
  // F4. Tt[(ru + r su), (rv + r sv)] = sum_b R_u[ru, su + du b] R_v[rv, sv + dv b]; Theta = G Tt
  for (int idx = tid; idx < Mr * Nc; idx += blockDim.x) {
    const int row = idx % Mr, col = idx / Mr;
    const int a = row % ru, su = row / ru, c = col % rv, sv = col / rv;
    double t = 0;
    for (int b = 0; b < chib; ++b) t += R[0][a + ru * (su + du * b)] * R[1][c + rv * (sv + dv * b)];
    Tt[idx] = t;
  }
  bsync();
  for (int idx = tid; idx < Mr * Nc; idx += blockDim.x) {
    const int row = idx % Mr, col = idx / Mr;
    const int a = row % ru, sup = row / ru, c = col % rv, svp = col / rv;
    double t = 0;
    for (int su = 0; su < du; ++su)
      for (int sv = 0; sv < dv; ++sv)
        t += G[(sup + du * svp) + dd * (su + du * sv)] * Tt[(a + ru * su) + Mr * (c + rv * sv)];
    Th[idx] = t;
  }
  bsync();
  // F5. thin SVD of Theta (one-sided Jacobi on the taller orientation)
  const bool tr = Mr < Nc;
  const int Ms = tr ? Nc : Mr, Ns = tr ? Mr : Nc;  // Ns = p
  const int p = Ns;
  for (int idx = tid; idx < Mr * Nc; idx += blockDim.x) {
    const int row = idx % Mr, col = idx / Mr;
    if (tr) As[col + Nc * row] = Th[idx];
    else As[idx] = Th[idx];
  }
  bsync();
  jacobi_svd(As, Ms, Ns, Vs, rot, red);
  bsync();
  for (int j = tid; j < Ns; j += blockDim.x) {
    double t = 0;
    for (int i = 0; i < Ms; ++i) t += As[i + Ms * j] * As[i + Ms * j];
    sbar[j] = sqrt(t);  // unsorted singular values (sbar reused below)
  }
  bsync();
  if (tid == 0) {
    for (int i = 0; i < Ns; ++i) order[i] = i;
    for (int i = 0; i < Ns; ++i)
      for (int j = i + 1; j < Ns; ++j)
        if (sbar[order[j]] > sbar[order[i]]) {
          const int t = order[i];
          order[i] = order[j];
          order[j] = t;
        }
    double tot = 0;
    for (int i = 0; i < Ns; ++i) tot += sbar[i] * sbar[i];
    int keep = Ns;
    if (prm.maxdim > 0 && keep > prm.maxdim) keep = prm.maxdim;
    double disc = 0;
    for (int i = keep; i < Ns; ++i) disc += sbar[order[i]] * sbar[order[i]];
    if (prm.cutoff >= 0 && tot > 0)
      while (keep > 1 && (disc + sbar[order[keep - 1]] * sbar[order[keep - 1]]) / tot <= prm.cutoff) {
        disc += sbar[order[keep - 1]] * sbar[order[keep - 1]];
        --keep;
      }
    int st = 0;
    if (keep != o.keep) st = 1;
    if (!isfinite(tot)) st = 2;
    for (int i = 0; i < keep; ++i)
      if (!(sbar[order[i]] > 0)) st = 3;
    status[blockIdx.x] = st;
    sh_keep = st ? -1 : keep;
  }
  __syncthreads();
  const int keep = sh_keep;
  if (keep < 0) return;
  // sorted thin factors: Um (Mr x p), Vm (Nc x p), svec (p)
  for (int idx = tid; idx < Mr * p; idx += blockDim.x) {
    const int i = idx % Mr, j = idx / Mr, oj = order[j];
    const double sj = sbar[oj];
    Um[idx] = tr ? Vs[i + Ns * oj] : (sj > 0 ? As[i + Ms * oj] / sj : 0.0);
  }
  for (int idx = tid; idx < Nc * p; idx += blockDim.x) {
    const int i = idx % Nc, j = idx / Nc, oj = order[j];
    const double sj = sbar[oj];
    Vm[idx] = tr ? (sj > 0 ? As[i + Ms * oj] / sj : 0.0) : Vs[i + Ns * oj];
  }
  bsync();
  for (int j = tid; j < p; j += blockDim.x) svec[j] = sbar[order[j]];
  bsync();
  // F6. F_u = U_k sqrt(S), F_v = V_k sqrt(S)
  for (int idx = tid; idx < Mr * keep; idx += blockDim.x) Fu[idx] = Um[idx] * sqrt(svec[idx / Mr]);
  for (int idx = tid; idx < Nc * keep; idx += blockDim.x) Fv[idx] = Vm[idx] * sqrt(svec[idx / Nc]);
  bsync();
  // F7. pre-gauge outputs Y_s[e, s' + d j] = sum_r W[e, r] F[r + r_s s', j] into T2
  const int rs[2] = {ru, rv};
  const double* Fs[2] = {Fu, Fv};
  for (int s = 0; s < 2; ++s) {
    const SideGeom& g = gs[s];
    const int r = rs[s];
    for (int idx = tid; idx < g.E * g.d * keep; idx += blockDim.x) {
      const int e = idx % g.E, c = idx / g.E, sp = c % g.d, j = c / g.d;
      double t = 0;
      for (int q = 0; q < r; ++q) t += W[s][e + (long long)g.E * q] * Fs[s][(q + r * sp) + (long long)rs[s] * g.d * j];
      T2[s][idx] = t;
    }
    bsync();
  }
  // bond signs from psi_u' (largest |entry| of every slice, first in natural order)
  if (tid < keep) {
    const SideGeom& g = gs[0];
    const int j = tid;
    double best = -1, val = 0;
    int best_off = 0x7fffffff;
    for (int e = 0; e < g.E; ++e)
      for (int s2 = 0; s2 < g.d; ++s2) {
        const double x = T2[0][e + (long long)g.E * (s2 + g.d * j)];
        const int off = nat_off(g, s2, e, j);
        if (fabs(x) > best || (fabs(x) == best && off < best_off)) {
          best = fabs(x);
          best_off = off;
          val = x;
        }
      }
    sgn[j] = val < 0 ? -1.0 : 1.0;
  }
  bsync();

  // ---- backward ------------------------------------------------------------
  // B1-B3. cotangents of the pre-gauge outputs: read, undo the normalisation, the signs
  for (int s = 0; s < 2; ++s) {
    const SideGeom& g = gs[s];
    SideGeom gk = g;
    gk.dims[g.j] = keep;
    const double* dsrc = dout + (s == 0 ? o.out_u : o.out_v);
    const int cols = g.d * keep;
    double n2 = 0, dot = 0;
    for (int idx = tid; idx < g.E * cols; idx += blockDim.x) {
      const int e = idx % g.E, c = idx / g.E;
      const double y = T2[s][idx] * sgn[c / g.d];
      const double yb = dsrc[nat_off(gk, c % g.d, e, c / g.d)];
      Yb[s][idx] = yb;
      n2 += y * y;
      dot += y * yb;
    }
    n2 = block_sum_d(n2, red);
    dot = block_sum_d(dot, red);
    const double nrm = sqrt(n2);
    const bool nz = prm.normalize && n2 > 0;
    for (int idx = tid; idx < g.E * cols; idx += blockDim.x) {
      const double y = T2[s][idx] * sgn[(idx / g.E) / g.d];
      double yb = Yb[s][idx];
      if (nz) yb = (yb - y * dot / n2) / nrm;  // pullback of y / ||y||
      Yb[s][idx] = yb * sgn[(idx / g.E) / g.d];
    }
    bsync();
  }
  // B4. Wbar (E x r) and Fbar (r d x keep)
  double* Fbs[2] = {Fub, Fvb};
  for (int s = 0; s < 2; ++s) {
    const SideGeom& g = gs[s];
    const int r = rs[s];
    for (int idx = tid; idx < g.E * r; idx += blockDim.x) {
      const int e = idx % g.E, q = idx / g.E;
      double t = 0;
      for (int j = 0; j < keep; ++j)
        for (int sp = 0; sp < g.d; ++sp)
          t += Yb[s][e + (long long)g.E * (sp + g.d * j)] * Fs[s][(q + r * sp) + (long long)r * g.d * j];
      T1[s][idx] = t;  // Wbar
    }
    for (int idx = tid; idx < r * g.d * keep; idx += blockDim.x) {
      const int row = idx % (r * g.d), j = idx / (r * g.d);
      const int q = row % r, sp = row / r;
      double t = 0;
      for (int e = 0; e < g.E; ++e) t += W[s][e + (long long)g.E * q] * Yb[s][e + (long long)g.E * (sp + g.d * j)];
      Fbs[s][idx] = t;
    }
    bsync();
  }
  // B5. Ubar = Fbar_u sqrt(S), Vbar = Fbar_v sqrt(S) (zero beyond keep), sbar
  for (int idx = tid; idx < Mr * p; idx += blockDim.x) {
    const int j = idx / Mr;
    M1[idx] = j < keep ? Fub[idx] * sqrt(svec[j]) : 0.0;  // Ubar
  }
  for (int idx = tid; idx < Nc * p; idx += blockDim.x) {
    const int j = idx / Nc;
    Q2[idx] = j < keep ? Fvb[idx] * sqrt(svec[j]) : 0.0;  // Vbar
  }
  for (int j = tid; j < p; j += blockDim.x) {
    double t = 0;
    if (j < keep) {
      for (int i = 0; i < Mr; ++i) t += Fub[i + Mr * j] * Um[i + Mr * j];
      for (int i = 0; i < Nc; ++i) t += Fvb[i + Nc * j] * Vm[i + Nc * j];
      t /= 2 * sqrt(svec[j]);
    }
    sbar[j] = t;
  }
  bsync();
  // B6. thin SVD pullback.  J = U^T Ubar, K = V^T Vbar (p x p) in Ttb, Thb
  gmm(p, p, Mr, Um, Mr, true, M1, Mr, false, Ttb, p);
  gmm(p, p, Nc, Vm, Nc, true, Q2, Nc, false, Thb, p);
  {
    const double smax = svec[0];
    const double eps2 = 1e-24 * smax * smax * smax * smax;
    // Inner (p x p) into Fu (free after B4), then M1' = U Inner + (Ubar - U J) S^-1
    for (int idx = tid; idx < p * p; idx += blockDim.x) {
      const int i = idx % p, j = idx / p;
      double v = i == j ? sbar[i] : 0.0;
      if (i != j) {
        const double dfr = svec[j] * svec[j] - svec[i] * svec[i];
        const double f = dfr / (dfr * dfr + eps2);
        v += f * (Ttb[i + p * j] - Ttb[j + p * i]) * svec[j] + svec[i] * f * (Thb[i + p * j] - Thb[j + p * i]);
      }
      Fu[idx] = v;
    }
    bsync();
  }
  // Fv <- U Inner (Mr x p); As <- U J (Mr x p)
  gmm(Mr, p, p, Um, Mr, false, Fu, p, false, Fv, Mr);
  gmm(Mr, p, p, Um, Mr, false, Ttb, p, false, As, Mr);
  for (int idx = tid; idx < Mr * p; idx += blockDim.x) {
    const int j = idx / Mr;
    const double inv = (j < keep && svec[j] > 0) ? 1.0 / svec[j] : 0.0;
    Fv[idx] += (M1[idx] - As[idx]) * inv;  // U Inner + (I - U U^T) Ubar S^-1
  }
  // Vs <- S^-1 (Vbar^T - K^T V^T) (p x Nc)
  bsync();
  gmm(p, Nc, p, Thb, p, true, Vm, Nc, true, Vs, p);  // K^T V^T
  for (int idx = tid; idx < p * Nc; idx += blockDim.x) {
    const int j = idx % p, c = idx / p;
    const double inv = (j < keep && svec[j] > 0) ? 1.0 / svec[j] : 0.0;
    Vs[idx] = (Q2[c + Nc * j] - Vs[idx]) * inv;
  }
  bsync();
  // Thetabar = Fv V^T + U Vs (Mr x Nc)
  gmm(Mr, Nc, p, Fv, Mr, false, Vm, Nc, true, Thb, Mr);
  gmm(Mr, Nc, p, Um, Mr, false, Vs, p, false, Thb, Mr, true);
  // B7. Ttbar and Gbar
  for (int idx = tid; idx < Mr * Nc; idx += blockDim.x) {
    const int row = idx % Mr, col = idx / Mr;
    const int a = row % ru, su = row / ru, c = col % rv, sv = col / rv;
    double t = 0;
    for (int sup = 0; sup < du; ++sup)
      for (int svp = 0; svp < dv; ++svp)
        t += G[(sup + du * svp) + dd * (su + du * sv)] * Thb[(a + ru * sup) + Mr * (c + rv * svp)];
    Ttb[idx] = t;
  }
  for (int idx = tid; idx < dd * dd; idx += blockDim.x) {
    const int i = idx % dd, jj = idx / dd;
    const int sup = i % du, svp = i / du, su = jj % du, sv = jj / du;
    double t = 0;
    for (int a = 0; a < ru; ++a)
      for (int c = 0; c < rv; ++c) t += Thb[(a + ru * sup) + Mr * (c + rv * svp)] * Tt[(a + ru * su) + Mr * (c + rv * sv)];
    dgate[o.dgate + idx] = t;
  }
  bsync();
  // B8. Rbar
  for (int idx = tid; idx < ru * nu; idx += blockDim.x) {
    const int a = idx % ru, cu = idx / ru, su = cu % du, b = cu / du;
    double t = 0;
    for (int c = 0; c < rv; ++c)
      for (int sv = 0; sv < dv; ++sv) t += Ttb[(a + ru * su) + Mr * (c + rv * sv)] * R[1][c + rv * (sv + dv * b)];
    Rb[0][idx] = t;
  }
  for (int idx = tid; idx < rv * nv; idx += blockDim.x) {
    const int c = idx % rv, cv = idx / rv, sv = cv % dv, b = cv / dv;
    double t = 0;
    for (int a = 0; a < ru; ++a)
      for (int su = 0; su < du; ++su) t += Ttb[(a + ru * su) + Mr * (c + rv * sv)] * R[0][a + ru * (su + du * b)];
    Rb[1][idx] = t;
  }
  bsync();
  // B9-B11 per side
  for (int s = 0; s < 2; ++s) {
    const SideGeom& g = gs[s];
    const int n = g.n;
    if (g.E >= n) {
      // Bbar = Pt^T Wbar (n x n); Ptbar = Wbar B (E x n) into Pb
      gmm(n, n, g.E, Pt[s], g.E, true, T1[s], g.E, false, Bb[s], n);
      gmm(g.E, n, n, T1[s], g.E, false, B[s], n, false, Pb[s], g.E);
      // Cbar = Z [(Z^T Rbar_s Z) o K1 + (Z^T Bbar_s Z) o K2] Z^T
      for (int idx = tid; idx < n * n; idx += blockDim.x) {
        const int i = idx % n, j = idx / n;
        Cb[s][idx] = 0.5 * (Rb[s][i + n * j] + Rb[s][j + n * i]);
        A[s][idx] = 0.5 * (Bb[s][i + n * j] + Bb[s][j + n * i]);
      }
      bsync();
      gmm(n, n, n, Z[s], n, true, Cb[s], n, false, T2[s], n);
      gmm(n, n, n, T2[s], n, false, Z[s], n, false, Cb[s], n);  // Z^T Rbar Z
      gmm(n, n, n, Z[s], n, true, A[s], n, false, T2[s], n);
      gmm(n, n, n, T2[s], n, false, Z[s], n, false, A[s], n);   // Z^T Bbar Z
      for (int idx = tid; idx < n * n; idx += blockDim.x) {
        const int i = idx % n, j = idx / n;
        const double si = Sg[s][i], sj = Sg[s][j];
        const double k1 = (si + sj) > 0 ? 1.0 / (si + sj) : 0.0;
        const double k2 = (si > 0 && sj > 0) ? -1.0 / (si * sj * (si + sj)) : 0.0;
        T2[s][idx] = Cb[s][idx] * k1 + A[s][idx] * k2;
      }
      bsync();
      gmm(n, n, n, Z[s], n, false, T2[s], n, false, A[s], n);
      gmm(n, n, n, A[s], n, false, Z[s], n, true, Cb[s], n);    // Cbar (symmetric)
      // Xbar = 2 X Cbar
      gmm(g.E, n, n, X[s], g.E, false, Cb[s], n, false, Xb[s], g.E);
      for (int i = tid; i < g.E * n; i += blockDim.x) Xb[s][i] *= 2.0;
      bsync();
    } else {
      for (int i = tid; i < g.E * n; i += blockDim.x) {
        Xb[s][i] = Rb[s][i];
        Pb[s][i] = 0.0;
      }
      bsync();
      // Sibar_a from Wbar (E x E) with W = (x) Si
      for (int pa = 0; pa < g.next; ++pa) {
        const int chi = g.dims[g.ext[pa]];
        double* Sib = env_p(s, pa, 4);
        // every (e, f) pair with e_pa = x, f_pa = y
        for (int idx = tid; idx < chi * chi; idx += blockDim.x) {
          const int x = idx / chi, y = idx % chi;
          double t = 0;
          for (int e = 0; e < g.E; ++e)
            for (int f = 0; f < g.E; ++f) {
              int re = e, rf = f;
              bool ok = true;
              double w = 1.0;
              for (int qa = 0; qa < g.next; ++qa) {
                const int c2 = g.dims[g.ext[qa]];
                const int ea = re % c2, fa = rf % c2;
                re /= c2;
                rf /= c2;
                if (qa == pa) ok = ok && ea == x && fa == y;
                else w *= env_p(s, qa, 2)[ea * c2 + fa];
              }
              if (ok) t += T1[s][e + (long long)g.E * f] * w;
            }
          Sib[x * chi + y] = t;
        }
        bsync();
      }
    }
    // Sbar_a = sum Xbar[e (e_a = x), c] (psi (x)_{others} S)[e (e_a = y), c]
    // (Xbar, Pb intact; T2, A, W, Pt are free)
    for (int pa = 0; pa < g.next; ++pa) {
      double* yc = T2[s];
      double* yo = A[s];
      for (int idx = tid; idx < g.E * n; idx += blockDim.x) {
        const int e = idx % g.E, c = idx / g.E;
        yc[idx] = psi[g.psi_off + nat_off(g, c % g.d, e, c / g.d)];
      }
      bsync();
      for (int qa = 0; qa < g.next; ++qa) {
        if (qa == pa) continue;
        ext_mp_g(g, qa, yc, yo, n, env_p(s, qa, 1));
        double* t = yc;
        yc = yo;
        yo = t;
      }
      int stride = 1;
      for (int i = 0; i < pa; ++i) stride *= g.dims[g.ext[i]];
      const int chi = g.dims[g.ext[pa]];
      double* Sb = env_p(s, pa, 3);
      for (int idx = tid; idx < chi * chi; idx += blockDim.x) {
        const int x = idx / chi, y = idx % chi;
        double t = 0;
        for (int c = 0; c < n; ++c)
          for (int rest = 0; rest < g.E / chi; ++rest) {
            const int lo = rest % stride, hi = rest / stride;
            t += Xb[s][(lo + stride * (x + chi * hi)) + (long long)g.E * c] *
                 yc[(lo + stride * (y + chi * hi)) + (long long)g.E * c];
          }
        Sb[x * chi + y] = t;
      }
      bsync();
    }
    // psibar = Xbar (x) S (+ Ptbar (x) P on a tall side), to the natural layout
    {
      const double* src = Xb[s];
      double* bufs2[2] = {T2[s], A[s]};
      int bi = 0;
      for (int pa = 0; pa < g.next; ++pa) {
        ext_mp_g(g, pa, src, bufs2[bi], n, env_p(s, pa, 1));
        src = bufs2[bi];
        bi ^= 1;
      }
      const double* srcp = nullptr;
      if (g.E >= n) {
        srcp = Pb[s];
        double* bufs3[2] = {W[s], Pt[s]};
        int bj = 0;
        for (int pa = 0; pa < g.next; ++pa) {
          ext_mp_g(g, pa, srcp, bufs3[bj], n, env_p(s, pa, 5));
          srcp = bufs3[bj];
          bj ^= 1;
        }
      }
      double* dst = din + (s == 0 ? o.in_u : o.in_v);
      for (int idx = tid; idx < g.E * n; idx += blockDim.x) {
        const int e = idx % g.E, c = idx / g.E;
        dst[nat_off(g, c % g.d, e, c / g.d)] = src[idx] + (srcp ? srcp[idx] : 0.0);
      }
      bsync();
    }
    // environment pullbacks into dmsg (accumulated; every message of a call belongs to one bond)
    {
      const int wv = tid >> 6;
      for (int pa = wv; pa < g.next; pa += (int)blockDim.x / 64) {
        const int a = g.ext[pa];
        const int chi = g.dims[a];
        double* Mb = env_p(s, pa, 5);
        env_pullback(env_p(s, pa, 0), env[s] + pa * VJ_ENV + 6 * 64, kmask[s][pa], chi, env_p(s, pa, 3),
                     env_p(s, pa, 4), env_p(s, pa, 2), Mb);
        __threadfence_block();
        const long long off = vds[vtx[s]].in_msg[a];
        for (int idx = tid & 63; idx < chi * chi; idx += 64) {
          const int i = idx % chi, j = idx / chi;  // message layout M[i + chi j]
          dmsg[off + idx] += Mb[i * chi + j];
        }
      }
      bsync();
    }
  }
}

}  // namespace

extern "C" int itnbp_apply_vjp(itnbp_net* net, int n, const int* edge_ids, const double* gates, int maxdim,
                               double cutoff, int normalize, const int* kept, const double* dpsi_out,
                               double* dpsi_in, double* dgates, double* dmsgs) {
  if (!net || n < 0 || (n > 0 && (!edge_ids || !gates || !kept || !dpsi_out || !dpsi_in || !dgates || !dmsgs)))
    return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  itnbp_ctx* ctx = net->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    std::vector<ApplyItem> items;
    long long goff = 0;
    ITNBP_RETCHK(prepare_apply(net, n, edge_ids, maxdim, cutoff, nullptr, 1, net->chi, false, items, goff));
    std::vector<VjpIo> io(n);
    long long out_tot = 0, in_tot = 0;
    for (int i = 0; i < n; ++i) {
      const ApplyItem& it = items[i];
      const int e = edge_ids[i];
      const int chib = net->chi[e];
      if (kept[i] < 1) return itnbp_fail(ctx, ITNBP_ERR_ARG, "apply_vjp: kept must be >= 1");
      for (int side = 0; side < 2; ++side) {
        const int x = side == 0 ? it.u : it.v;
        const long long D = net->vdesc[x].D;
        const long long E = D / ((long long)net->phys[x] * chib);
        if (E * net->phys[x] * kept[i] > AP_XY)
          return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "apply_vjp: the new site tensor is too large");
        const long long Dn = D / chib * kept[i];
        if (side == 0) {
          io[i].out_u = out_tot;
          io[i].in_u = in_tot;
        } else {
          io[i].out_v = out_tot;
          io[i].in_v = in_tot;
        }
        out_tot += Dn;
        in_tot += D;
      }
      io[i].dgate = it.gate;
      io[i].keep = kept[i];
      io[i].pad = 0;
    }
    DevBuf b_items, b_io, b_gates, b_out, b_in, b_dg, b_dm, b_work, b_st;
    struct Guard {
      DevBuf* b[9];
      ~Guard() {
        for (DevBuf* x : b) free_buf(*x);
      }
    } guard{{&b_items, &b_io, &b_gates, &b_out, &b_in, &b_dg, &b_dm, &b_work, &b_st}};
    ITNBP_RETCHK(upload(ctx, b_items, items.data(), n * sizeof(ApplyItem)));
    ITNBP_RETCHK(upload(ctx, b_io, io.data(), n * sizeof(VjpIo)));
    ITNBP_RETCHK(upload(ctx, b_gates, gates, goff * 8));
    ITNBP_RETCHK(upload(ctx, b_out, dpsi_out, out_tot * 8));
    ITNBP_RETCHK(upload(ctx, b_dm, dmsgs, net->msg_total * 8));
    ITNBP_RETCHK(ensure_buf(ctx, b_in, in_tot * 8));
    ITNBP_RETCHK(ensure_buf(ctx, b_dg, goff * 8));
    ITNBP_RETCHK(ensure_buf(ctx, b_work, (size_t)n * VJ_GATE * 8));
    ITNBP_RETCHK(ensure_buf(ctx, b_st, n * sizeof(int)));
    ApplyParams prm;
    prm.maxdim = maxdim;
    prm.cutoff = cutoff;
    prm.normalize = normalize;
    prm.eig_cutoff = 10 * 2.220446049250313e-16;
    prm.max_keep = 1;
    const int pm0 = prof_mark(ctx);
    hipLaunchKernelGGL(k_apply_vjp, dim3(n), dim3(256), 0, ctx->stream, (const ApplyItem*)b_items.p,
                       (const VjpIo*)b_io.p, net->d_vdesc, (const double*)net->d_psi, net->d_msg[net->cur],
                       (const double*)b_gates.p, prm, (const double*)b_out.p, (double*)b_in.p, (double*)b_dg.p,
                       (double*)b_dm.p, (double*)b_work.p, (int*)b_st.p);
    ITNBP_HIPCHK(ctx, hipGetLastError());
    prof_add(ctx, "apply_vjp", pm0, prof_mark(ctx));
    std::vector<int> st(n);
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(st.data(), b_st.p, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(dpsi_in, b_in.p, in_tot * 8, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(dgates, b_dg.p, goff * 8, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(dmsgs, b_dm.p, net->msg_total * 8, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
      if (st[i] == 1)
        return itnbp_fail(ctx, ITNBP_ERR_ARG, "apply_vjp: kept differs from the rank the truncation keeps");
      if (st[i] == 2) return itnbp_fail(ctx, ITNBP_ERR_NONFINITE, "apply_vjp: non-finite two-site tensor");
      if (st[i] == 3)
        return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "apply_vjp: a kept singular value is zero (no derivative)");
    }
    return ITNBP_OK;
  });
}
