// Per-vertex eigenbasis message update for small bond dimensions (every
// chi <= 16, the site tensor resident in LDS).  One workgroup computes all k
// outgoing messages of a vertex in a Jacobi sweep
// (updated_message, src/caches/abstractbeliefpropagationcache.jl:225-239):
//
//   1. psi_v -> LDS;
//   2. every incoming message M_c (symmetric) -> M_c = U_c diag(l_c) U_c^T by a
//      wave-level cyclic Jacobi (the four waves take one bond each);
//   3. psi_v rotated in place into the eigenbasis of every bond (fibre by fibre);
//   4. per bond c: T'_c[r, r'] = sum over the fibres along c of
//      prod_{e != c} l_e[r_e] Y[.., r, ..] Y[.., r', ..];
//   5. T_c = U_c T'_c U_c^T, symmetrised, normalised, message_diff / residual.
//
// That is d prod(chi) (k chi + k chi) multiply-adds per vertex for all k
// messages, versus k (k - 1) chi d prod(chi) for k independent contractions, and
// psi is read from HBM once per vertex instead of once per message.  The
// absorbed form is recovered exactly for any orthogonal U_c (eigenvalues of
// either sign), so the result is the reference's update for symmetric messages.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "itnbp_ops.hpp"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int SV_MAXCHI = 16;
constexpr int SV_LD = SV_MAXCHI + 1;
constexpr int SV_THREADS = 256;

struct SmallTask {
  int v;
  int inst;                   // batched update: instance of the vertex (0 otherwise)
  long long out[ITNBP_MAXK];  // outgoing message on bond axis a + 1
  int slot[ITNBP_MAXK];
  long long yoff;             // k_vertex_big: offset of the site's scratch tensor
  int useed[ITNBP_MAXK];      // id of the incoming message on bond axis a + 1 (eigenvector seed slot)
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
}

// Symmetric n x n (n <= 16) eigendecomposition in LDS by one wave: A (ld
// SV_LD) is overwritten by diag(l), V receives the eigenvectors as columns.
// With init_v = false, V holds an orthogonal seed onto which the rotations
// accumulate.
__device__ void wave_jacobi_small(double* A, double* V, int n, double* rc, double* rs, int* rp, int* rq,
                                  bool init_v = true) {
  const int l = threadIdx.x & 63;
  const int ne = n + (n & 1), np = ne / 2;
  if (init_v)
    for (int o = l; o < n * n; o += 64) {
      const int i = o / n, j = o % n;
      V[i * SV_LD + j] = (i == j) ? 1.0 : 0.0;
    }
  wave_sync();
  double prev = INFINITY;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0, dg = 0;
    for (int o = l; o < n * n; o += 64) {
      const int i = o / n, j = o % n;
      const double a = A[i * SV_LD + j];
      if (i != j) off += a * a;
      else dg += a * a;
    }
    for (int m = 32; m > 0; m >>= 1) {
      off += __shfl_xor(off, m);
      dg += __shfl_xor(dg, m);
    }
    // converged (off-diagonal mass at rounding level) or no longer decreasing
    if (off <= 1e-28 * dg || off == 0.0 || off >= prev) break;
    prev = off;
    for (int st = 0; st < ne - 1; ++st) {
      if (l < np) {
        auto at = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + st) % (ne - 1); };
        int p = at(l), q = at(ne - 1 - l);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1, s = 0;
        if (q < n) {
          const double apq = A[p * SV_LD + q];
          if (apq != 0.0) {
            const double zeta = (A[q * SV_LD + q] - A[p * SV_LD + p]) / (2 * apq);
            const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
            c = 1 / sqrt(1 + t * t);
            s = c * t;
          }
        }
        rc[l] = c;
        rs[l] = s;
        rp[l] = p;
        rq[l] = q;
      }
      wave_sync();
      for (int o = l; o < np * n; o += 64) {  // rows: A <- J^T A
        const int i = o / n, col = o % n;
        const int p = rp[i], q = rq[i];
        if (q >= n) continue;
        const double c = rc[i], s = rs[i];
        const double ap = A[p * SV_LD + col], aq = A[q * SV_LD + col];
        A[p * SV_LD + col] = c * ap - s * aq;
        A[q * SV_LD + col] = s * ap + c * aq;
      }
      wave_sync();
      for (int o = l; o < np * n; o += 64) {  // columns: A <- A J, V <- V J
        const int i = o / n, row = o % n;
        const int p = rp[i], q = rq[i];
        if (q >= n) continue;
        const double c = rc[i], s = rs[i];
        const double ap = A[row * SV_LD + p], aq = A[row * SV_LD + q];
        A[row * SV_LD + p] = c * ap - s * aq;
        A[row * SV_LD + q] = s * ap + c * aq;
        const double vp = V[row * SV_LD + p], vq = V[row * SV_LD + q];
        V[row * SV_LD + p] = c * vp - s * vq;
        V[row * SV_LD + q] = s * vp + c * vq;
      }
      wave_sync();
    }
  }
}

constexpr int SEED_LD = SV_MAXCHI;  // row stride of a stored eigenvector seed

// n x n product by one wave into registers (at most 4 elements per lane for
// n <= 16): C[i][j] = sum_k X(i, k) Y(k, j), X read transposed when tx
__device__ inline void wave_mm(const double* X, bool tx, const double* Y, int n, double (&c)[4]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int o = l + 64 * q;
    c[q] = 0.0;
    if (o < n * n) {
      const int i = o / n, j = o % n;
      double s = 0.0;
      for (int kk = 0; kk < n; ++kk) s += (tx ? X[kk * SV_LD + i] : X[i * SV_LD + kk]) * Y[kk * SV_LD + j];
      c[q] = s;
    }
  }
}

// Eigendecomposition of the symmetrised message M (n x n, column-major) on one
// bond by one wave: eigenvectors to U (columns, ld SV_LD), eigenvalues to lamv,
// A (ld SV_LD) as work.  With warm, the previous sweep's eigenvectors U0 (in
// seed, ld SEED_LD) start it: Jacobi diagonalises the symmetrised U0^T S U0 and
// its rotations accumulate onto U0; near a BP fixed point that matrix is
// nearly diagonal and takes one or two sweeps instead of a cold start's six to
// ten.  The eigenvectors are stored back to seed for the next sweep.
template <typename TM>
__device__ void bond_eigh(const TM* M, int n, double* A, double* U, double* lamv, double* seed, int warm,
                          double* rc, double* rs, int* rp, int* rq) {
  const int l = threadIdx.x & 63;
  for (int o = l; o < n * n; o += 64) {
    const int i = o / n, j = o % n;
    A[i * SV_LD + j] = 0.5 * ((double)M[i + n * j] + (double)M[j + n * i]);
    if (warm) U[i * SV_LD + j] = seed[i * SEED_LD + j];
  }
  wave_sync();
  if (warm) {
    double c[4];
    wave_mm(A, false, U, n, c);  // S U0
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (l + 64 * q < n * n) A[((l + 64 * q) / n) * SV_LD + (l + 64 * q) % n] = c[q];
    wave_sync();
    wave_mm(U, true, A, n, c);  // U0^T (S U0)
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (l + 64 * q < n * n) A[((l + 64 * q) / n) * SV_LD + (l + 64 * q) % n] = c[q];
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // symmetrise
      const int o = l + 64 * q;
      if (o < n * n) c[q] = 0.5 * (A[(o / n) * SV_LD + o % n] + A[(o % n) * SV_LD + o / n]);
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (l + 64 * q < n * n) A[((l + 64 * q) / n) * SV_LD + (l + 64 * q) % n] = c[q];
    wave_sync();
  }
  wave_jacobi_small(A, U, n, rc, rs, rp, rq, !warm);
  if (l < n) lamv[l] = A[l * SV_LD + l];
  for (int o = l; o < n * n; o += 64) seed[(o / n) * SEED_LD + o % n] = U[(o / n) * SV_LD + o % n];
  wave_sync();
}

__device__ inline double block_sum(double v, double* red) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < SV_THREADS / 64; ++i) s += red[i];
  return s;
}

// LDS carve-up (bytes): fixed part, the fibre tables (FMAX weights and base
// offsets), then Y[D] in TA
constexpr size_t SV_FIXED = (size_t)ITNBP_MAXK * SV_LD * SV_LD * 8   // U_c
                            + 4 * SV_LD * SV_LD * 8                   // Jacobi work, one per wave
                            + (size_t)ITNBP_MAXK * SV_MAXCHI * 8      // l_c
                            + 4 * 8 * (8 + 8 + 4 + 4)                 // rotations
                            + SV_THREADS * 8                          // group partials of T'
                            + 2 * SV_LD * SV_LD * 8                   // T', U T'
                            + 64;                                     // block reduction

// One site's update (the task tk) by the whole workgroup; k_vertex_eig runs
// one task per workgroup, k_vertex_eig_queue a persistent grid that takes
// tasks from a queue.
template <typename TS, typename TA, typename TM>
__device__ __forceinline__ void vertex_eig_task(const VDesc* __restrict__ vdesc, const SmallTask* __restrict__ tk,
                                                const TS* __restrict__ psi, const TM* msg_in, TM* msg_out,
                                                MsgStat* __restrict__ stats, int normalize, int fmax,
                                                double* __restrict__ seeds, int warm,
                                                const int* __restrict__ active,
                                                const long long* __restrict__ sweep_ctr) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* Us = reinterpret_cast<double*>(smem);
  double* As = Us + ITNBP_MAXK * SV_LD * SV_LD;
  double* lam = As + 4 * SV_LD * SV_LD;
  double* rc = lam + ITNBP_MAXK * SV_MAXCHI;
  double* rs = rc + 32;
  int* rp = reinterpret_cast<int*>(rs + 32);
  int* rq = rp + 32;
  double* gpart = reinterpret_cast<double*>(rq + 32);
  double* Tp = gpart + SV_THREADS;
  double* X = Tp + SV_LD * SV_LD;
  double* red = X + SV_LD * SV_LD;
  double* fw = reinterpret_cast<double*>(smem + SV_FIXED);
  int* fb = reinterpret_cast<int*>(fw + fmax);
  TA* Y = reinterpret_cast<TA*>(smem + SV_FIXED + (((size_t)fmax * 12 + 15) & ~(size_t)15));

  __shared__ int dims[ITNBP_MAXR], strides[ITNBP_MAXR];
  // batched update: a finished instance's messages were carried over already
  if (active && !active[tk->inst]) return;
  if (sweep_ctr) warm = (*sweep_ctr % 16) != 0;
  const VDesc* V = vdesc + tk->v;
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
  const int k = V->k, D = V->D;
  if (tid == 0) {
    int s = 1;
    for (int a = 0; a <= k; ++a) {
      dims[a] = V->dims[a];
      strides[a] = s;
      s *= dims[a];
    }
  }
  __syncthreads();
  // 1. psi -> LDS
  const TS* src = psi + V->psi_off;
  for (int i = tid; i < D; i += SV_THREADS) Y[i] = load_as<TA>(src, i);
  // 2. eigendecompositions, one bond per wave
  for (int a = wv; a < k; a += 4)
    bond_eigh(msg_in + V->in_msg[a + 1], dims[a + 1], As + wv * SV_LD * SV_LD, Us + a * SV_LD * SV_LD,
              lam + a * SV_MAXCHI, seeds + (long long)tk->useed[a] * SEED_LD * SEED_LD, warm, rc + 8 * wv,
              rs + 8 * wv, rp + 8 * wv, rq + 8 * wv);
  __syncthreads();
  // 3. rotate every bond: Y <- Y x_c U_c (in place, one fibre per thread)
  for (int a = 0; a < k; ++a) {
    const int n = dims[a + 1], st = strides[a + 1];
    const int F = D / n;
    const double* U = Us + a * SV_LD * SV_LD;
    if (n <= 8) {
      // U in registers (zero-padded to 8 x 8), read once per bond
      TA u[8][8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int r = 0; r < 8; ++r) u[j][r] = (j < n && r < n) ? (TA)U[j * SV_LD + r] : (TA)0;
      for (int f = tid; f < F; f += SV_THREADS) {
        const int base = (f % st) + (f / st) * st * n;
        TA x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = j < n ? Y[base + j * st] : (TA)0;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          TA acc = 0;
#pragma unroll
          for (int j = 0; j < 8; ++j) acc += u[j][r] * x[j];
          if (r < n) Y[base + r * st] = acc;
        }
      }
      __syncthreads();
      continue;
    }
    for (int f = tid; f < F; f += SV_THREADS) {
      const int base = (f % st) + (f / st) * st * n;
      TA x[SV_MAXCHI];
#pragma unroll
      for (int j = 0; j < SV_MAXCHI; ++j) x[j] = j < n ? Y[base + j * st] : (TA)0;
      for (int r = 0; r < n; ++r) {
        TA acc = 0;
#pragma unroll
        for (int j = 0; j < SV_MAXCHI; ++j)
          if (j < n) acc += (TA)U[j * SV_LD + r] * x[j];
        Y[base + r * st] = acc;
      }
    }
    __syncthreads();
  }
  // 4-5. per bond: weighted Gram, rotate back, normalise, statistics
  for (int a = 0; a < k; ++a) {
    const int n = dims[a + 1], st = strides[a + 1], nn = n * n;
    const int F = D / n;
    // fibre table: base offset and weight prod_{b != a} l_b of every fibre along a
    for (int f = tid; f < F; f += SV_THREADS) {
      const int base = (f % st) + (f / st) * st * n;
      double w = 1.0;
      for (int b = 0; b < k; ++b)
        if (b != a) w *= lam[b * SV_MAXCHI + (base / strides[b + 1]) % dims[b + 1]];
      fb[f] = base;
      fw[f] = w;
    }
    __syncthreads();
    if (n <= 8) {
      // lane (r, gq) of wave wv: row r = l & 7 of T' over the fibres of group
      // wv * 8 + gq (gq = l >> 3), all n columns at once; fibre groups summed with
      // shuffles inside the wave, then across the four waves
      const int r = l & 7, gg = wv * 8 + (l >> 3);
      double acc8[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) acc8[j] = 0;
      if (r < n) {
        // four fibres per step, their loads issued together
        int f = gg;
        for (; f + 96 < F; f += 128) {
          int b[4];
          double wy[4], y[4][8];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            b[i] = fb[f + 32 * i];
            wy[i] = fw[f + 32 * i];
          }
#pragma unroll
          for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 8; ++j) y[i][j] = j < n ? (double)Y[b[i] + j * st] : 0.0;
            wy[i] *= (double)Y[b[i] + r * st];
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) acc8[j] += (wy[0] * y[0][j] + wy[1] * y[1][j]) + (wy[2] * y[2][j] + wy[3] * y[3][j]);
        }
        for (; f < F; f += 32) {
          const int b = fb[f];
          const double wy = fw[f] * (double)Y[b + r * st];
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (j < n) acc8[j] += wy * (double)Y[b + j * st];
        }
      }
#pragma unroll
      for (int j = 0; j < 8; ++j)
        for (int m = 8; m < 64; m <<= 1) acc8[j] += __shfl_xor(acc8[j], m);
      if (l < 8) {
#pragma unroll
        for (int j = 0; j < 8; ++j) gpart[wv * 64 + l * 8 + j] = acc8[j];
      }
      __syncthreads();
      if (tid < nn) {
        const int rr = tid % n, r2 = tid / n;
        Tp[rr * SV_LD + r2] = (gpart[rr * 8 + r2] + gpart[64 + rr * 8 + r2]) +
                              (gpart[128 + rr * 8 + r2] + gpart[192 + rr * 8 + r2]);
      }
      __syncthreads();
    }
    // T' = sum_f w_f y_f y_f^T on the FP64 matrix cores: v_mfma_f64_16x16x4f64
    // with A[i][k] = w_f Y[f, i] and B[k][j] = Y[f, j] for the four fibres
    // f = 4 b + k of a fibre block b (lane i + 16 k loads Y[f, i] once; rows and
    // columns >= n are zero); wave wv takes the blocks b = wv (mod 4), two
    // accumulators in turn so that consecutive MFMAs do not wait on each other.
    // D[4 v + l / 16][l % 16] is value v of lane l; the four waves' partial
    // Grams meet in the Jacobi work area (idle after step 2) and are summed in
    // a fixed order.  (For n <= 8 the VALU form above is faster: three
    // quarters of the 16 x 16 tile would be padding.)
    if (n > 8) {
      const int li = l & 15, lk = l >> 4;
      f64x4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
      const int nblk = (F + 3) >> 2;
      int b4 = wv;
      for (; b4 + 4 < nblk; b4 += 8) {
        const int f0 = 4 * b4 + lk, f1 = f0 + 16;
        double y0 = 0, y1 = 0, w0 = 0, w1 = 0;
        if (li < n) {
          if (f0 < F) {
            y0 = (double)Y[fb[f0] + li * st];
            w0 = fw[f0];
          }
          if (f1 < F) {
            y1 = (double)Y[fb[f1] + li * st];
            w1 = fw[f1];
          }
        }
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(w0 * y0, y0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(w1 * y1, y1, acc1, 0, 0, 0);
      }
      for (; b4 < nblk; b4 += 4) {
        const int f0 = 4 * b4 + lk;
        double y0 = 0, w0 = 0;
        if (li < n && f0 < F) {
          y0 = (double)Y[fb[f0] + li * st];
          w0 = fw[f0];
        }
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(w0 * y0, y0, acc0, 0, 0, 0);
      }
      double* part = As;  // 4 x 256 doubles
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 4 * v + lk;
        if (row < n && li < n) part[wv * 256 + row * 16 + li] = acc0[v] + acc1[v];
      }
      __syncthreads();
      if (tid < nn) {
        const int rr = tid % n, r2 = tid / n;
        Tp[rr * SV_LD + r2] = (part[rr * 16 + r2] + part[256 + rr * 16 + r2]) +
                              (part[512 + rr * 16 + r2] + part[768 + rr * 16 + r2]);
      }
      __syncthreads();
    }
    const double* U = Us + a * SV_LD * SV_LD;
    if (tid < nn) {
      const int b = tid % n, rr = tid / n;
      double s = 0;
      for (int r = 0; r < n; ++r) s += U[b * SV_LD + r] * Tp[r * SV_LD + rr];
      X[b * SV_LD + rr] = s;  // (U T')[b][r']
    }
    __syncthreads();
    double tv = 0;
    if (tid < nn) {
      const int i = tid % n, j = tid / n;  // message layout M[i + n j]
      double sij = 0, sji = 0;
      for (int r = 0; r < n; ++r) {
        sij += X[i * SV_LD + r] * U[j * SV_LD + r];
        sji += X[j * SV_LD + r] * U[i * SV_LD + r];
      }
      tv = 0.5 * (sij + sji);
    }
    const double n2 = block_sum(tv * tv, red);
    const double nrm = sqrt(n2);
    const double sc = (normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
    const long long off = tk->out[a];
    double dot = 0, o2 = 0, d2 = 0, bad = 0;
    if (tid < nn) {
      const double nv = tv * sc;
      const double ov = (double)msg_in[off + tid];
      if (!isfinite(nv)) bad = 1;
      dot = ov * nv;
      o2 = ov * ov;
      d2 = (nv - ov) * (nv - ov);
      msg_out[off + tid] = (TM)nv;
    }
    dot = block_sum(dot, red);
    o2 = block_sum(o2, red);
    d2 = block_sum(d2, red);
    bad = block_sum(bad, red);
    const int slot = tk->slot[a];
    if (tid == 0 && slot >= 0) {
      const double nn2 = n2 * sc * sc;
      MsgStat stt;
      stt.diff = 1.0 - (dot * dot) / (o2 * nn2);
      stt.res2 = nn2 > 0 ? d2 / nn2 : 0.0;
      stt.norm = nrm;
      stt.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
      stats[slot] = stt;
    }
    __syncthreads();
  }
}

template <typename TS, typename TA, typename TM>
__global__ __launch_bounds__(SV_THREADS) void k_vertex_eig(const VDesc* __restrict__ vdesc,
                                                          const SmallTask* __restrict__ tasks,
                                                          const TS* __restrict__ psi, const TM* msg_in,
                                                          TM* msg_out, MsgStat* __restrict__ stats,
                                                          int normalize, int fmax, double* __restrict__ seeds,
                                                          int warm, const int* __restrict__ active,
                                                          const long long* __restrict__ sweep_ctr) {
  vertex_eig_task<TS, TA, TM>(vdesc, tasks + blockIdx.x, psi, msg_in, msg_out, stats, normalize, fmax, seeds, warm,
                              active, sweep_ctr);
}

// Persistent form: a grid of resident workgroups takes the (cost-sorted) tasks
// in order from a device counter that k_queue_reset clears before the launch;
// every workgroup leaves once the counter has passed ntasks, so the grid
// drains without any wait on another workgroup.
template <typename TS, typename TA, typename TM>
__global__ __launch_bounds__(SV_THREADS) void k_vertex_eig_queue(const VDesc* __restrict__ vdesc,
                                                                const SmallTask* __restrict__ tasks, int ntasks,
                                                                int* __restrict__ queue, const TS* __restrict__ psi,
                                                                const TM* msg_in, TM* msg_out,
                                                                MsgStat* __restrict__ stats, int normalize, int fmax,
                                                                double* __restrict__ seeds, int warm,
                                                                const int* __restrict__ active,
                                                                const long long* __restrict__ sweep_ctr) {
  __shared__ int next;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(queue, 1);
    __syncthreads();
    const int t = next;
    __syncthreads();  // every thread has read next before thread 0 takes the following task
    if (t >= ntasks) return;
    vertex_eig_task<TS, TA, TM>(vdesc, tasks + t, psi, msg_in, msg_out, stats, normalize, fmax, seeds, warm, active,
                                sweep_ctr);
    __syncthreads();  // the task's last LDS reads come before the next task's writes
  }
}

// clears the task counter of k_vertex_eig_queue (a kernel, so that it can sit
// in a captured graph; one lane writes)
__global__ void k_queue_reset(int* __restrict__ queue) {
  if (threadIdx.x == 0) queue[0] = 0;
}

// The same update for sites whose tensor does not fit in LDS (every chi <= 16,
// e.g. chi = 12 or 16 at degree 4): psi is copied into a scratch tensor Y in
// global memory (L2 / Infinity Cache resident while the workgroup works on it),
// rotated there in place fibre by fibre, and every weighted Gram streams Y
// through LDS in tiles of FT fibres (base offsets and weights computed by the
// loader).  Steps 2 and 5 are those of k_vertex_eig.  The tile size is a
// template parameter; the batched update's autotuner times the VB_TILES
// variants and keeps the fastest, VB_FT otherwise.
constexpr int VB_FT = 192;
constexpr int VB_TILES[3] = {160, 192, 256};
// fixed LDS of k_vertex_big: that of k_vertex_eig without the Jacobi work
// areas, which alias the Gram tile (idle while the messages are decomposed), so
// that four FP32 workgroups share a CU
constexpr size_t VB_FIXED = SV_FIXED - 4 * SV_LD * SV_LD * 8;
static_assert(160 * SV_LD * sizeof(float) >= 4 * SV_LD * SV_LD * 8, "Jacobi work must fit in the tile area");

template <typename TA>
constexpr size_t vb_lds(int ft) {
  return VB_FIXED + (size_t)ft * 8 + (size_t)ft * SV_LD * sizeof(TA);
}

template <typename TS, typename TA, typename TM, int FT>
__global__ __launch_bounds__(SV_THREADS) void k_vertex_big(const VDesc* __restrict__ vdesc,
                                                          const SmallTask* __restrict__ tasks,
                                                          const TS* __restrict__ psi, const TM* msg_in,
                                                          TM* msg_out, MsgStat* __restrict__ stats, int normalize,
                                                          TA* __restrict__ scratch, double* __restrict__ seeds,
                                                          int warm, const int* __restrict__ active,
                                                          const long long* __restrict__ sweep_ctr) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* Us = reinterpret_cast<double*>(smem);
  double* lam = Us + ITNBP_MAXK * SV_LD * SV_LD;
  double* rc = lam + ITNBP_MAXK * SV_MAXCHI;
  double* rs = rc + 32;
  int* rp = reinterpret_cast<int*>(rs + 32);
  int* rq = rp + 32;
  double* gpart = reinterpret_cast<double*>(rq + 32);
  double* Tp = gpart + SV_THREADS;
  double* X = Tp + SV_LD * SV_LD;
  double* red = X + SV_LD * SV_LD;
  double* tw = reinterpret_cast<double*>(smem + VB_FIXED);      // tile weights [FT]
  TA* tv8 = reinterpret_cast<TA*>(tw + FT);                    // tile values [FT][SV_LD]
  double* As = reinterpret_cast<double*>(tv8);                  // Jacobi work (step 2 only)

  __shared__ int dims[ITNBP_MAXR], strides[ITNBP_MAXR];
  const SmallTask* tk = tasks + blockIdx.x;
  // batched update: a finished instance's messages were carried over already
  if (active && !active[tk->inst]) return;
  if (sweep_ctr) warm = (*sweep_ctr % 16) != 0;
  const VDesc* V = vdesc + tk->v;
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
  const int k = V->k, D = V->D;
  if (tid == 0) {
    int s = 1;
    for (int a = 0; a <= k; ++a) {
      dims[a] = V->dims[a];
      strides[a] = s;
      s *= dims[a];
    }
  }
  __syncthreads();
  TA* Y = scratch + tk->yoff;
  const TS* src = psi + V->psi_off;
  // 2. eigendecompositions, one bond per wave
  for (int a = wv; a < k; a += 4)
    bond_eigh(msg_in + V->in_msg[a + 1], dims[a + 1], As + wv * SV_LD * SV_LD, Us + a * SV_LD * SV_LD,
              lam + a * SV_MAXCHI, seeds + (long long)tk->useed[a] * SEED_LD * SEED_LD, warm, rc + 8 * wv,
              rs + 8 * wv, rp + 8 * wv, rq + 8 * wv);
  __threadfence_block();
  __syncthreads();
  // 3. rotate every bond, NFB fibres per thread in flight (their loads issued
  // together, each U element read once for all of them); bond 0 reads psi and
  // writes Y, the later bonds rotate Y in place
  constexpr int NFB = sizeof(TA) == 4 ? 4 : 2;
  for (int a = 0; a < k; ++a) {
    const int n = dims[a + 1], st = strides[a + 1];
    const int F = D / n;
    const double* U = Us + a * SV_LD * SV_LD;
    for (int f0 = tid; f0 < F; f0 += SV_THREADS * NFB) {
      int base[NFB];
      TA x[NFB][SV_MAXCHI];
#pragma unroll
      for (int i = 0; i < NFB; ++i) {
        const int f = f0 + i * SV_THREADS;
        base[i] = f < F ? (f % st) + (f / st) * st * n : -1;
#pragma unroll
        for (int j = 0; j < SV_MAXCHI; ++j)
          x[i][j] = (base[i] >= 0 && j < n) ? (a == 0 ? load_as<TA>(src, base[i] + j * st) : Y[base[i] + j * st])
                                             : (TA)0;
      }
      for (int r = 0; r < n; ++r) {
        TA acc[NFB];
#pragma unroll
        for (int i = 0; i < NFB; ++i) acc[i] = 0;
#pragma unroll
        for (int j = 0; j < SV_MAXCHI; ++j)
          if (j < n) {
            const TA u = (TA)U[j * SV_LD + r];
#pragma unroll
            for (int i = 0; i < NFB; ++i) acc[i] += u * x[i][j];
          }
#pragma unroll
        for (int i = 0; i < NFB; ++i)
          if (base[i] >= 0) Y[base[i] + r * st] = acc[i];
      }
    }
    __threadfence_block();
    __syncthreads();
  }
  // 4-5. per bond: weighted Gram over tiles of fibres, rotate back, normalise, statistics
  for (int a = 0; a < k; ++a) {
    const int n = dims[a + 1], st = strides[a + 1], nn = n * n;
    const int F = D / n;
    // lane (rl, grp): row rl = l & 15 of T' over the tile fibres j = grp mod 16
    // (grp = (l >> 4) + 4 wv), every column at once
    const int rl = l & 15, grp = (l >> 4) + 4 * wv;
    double acc16[SV_MAXCHI];
#pragma unroll
    for (int x = 0; x < SV_MAXCHI; ++x) acc16[x] = 0;
    // loader: thread j of a tile -> fibre f0 + j (values and weight), fetched
    // into registers one tile ahead so that the loads overlap the Gram of the
    // current tile
    TA pv[SV_MAXCHI];
    double pw = 0;
    auto fetch = [&](int f0) {
      if (tid < min(FT, F - f0)) {
        const int f = f0 + tid;
        const int base = (f % st) + (f / st) * st * n;
        double w = 1.0;
        for (int b = 0; b < k; ++b)
          if (b != a) w *= lam[b * SV_MAXCHI + (base / strides[b + 1]) % dims[b + 1]];
        pw = w;
#pragma unroll
        for (int x = 0; x < SV_MAXCHI; ++x) pv[x] = x < n ? Y[base + x * st] : (TA)0;
      }
    };
    fetch(0);
    for (int f0 = 0; f0 < F; f0 += FT) {
      const int nf = min(FT, F - f0);
      if (tid < nf) {
        tw[tid] = pw;
#pragma unroll
        for (int x = 0; x < SV_MAXCHI; ++x)
          if (x < n) tv8[tid * SV_LD + x] = pv[x];
      }
      __syncthreads();
      if (f0 + FT < F) fetch(f0 + FT);
      if (rl < n) {
#pragma unroll 2
        for (int j = grp; j < nf; j += 16) {
          const double wy = tw[j] * (double)tv8[j * SV_LD + rl];
#pragma unroll
          for (int x = 0; x < SV_MAXCHI; ++x)
            if (x < n) acc16[x] += wy * (double)tv8[j * SV_LD + x];
        }
      }
      __syncthreads();
    }
    // the 16 fibre groups: four per wave by shuffles, then the four waves in LDS
    // (the tile area, free after the loop: 4 x 256 doubles)
#pragma unroll
    for (int x = 0; x < SV_MAXCHI; ++x)
      for (int m = 16; m < 64; m <<= 1) acc16[x] += __shfl_xor(acc16[x], m);
    double* red4 = reinterpret_cast<double*>(tv8);
    if (l < 16) {
#pragma unroll
      for (int x = 0; x < SV_MAXCHI; ++x) red4[wv * 256 + l * 16 + x] = acc16[x];
    }
    __syncthreads();
    if (tid < nn) {
      const int rr = tid % n, r2 = tid / n;
      Tp[rr * SV_LD + r2] = (red4[rr * 16 + r2] + red4[256 + rr * 16 + r2]) +
                            (red4[512 + rr * 16 + r2] + red4[768 + rr * 16 + r2]);  // T'[r][r']
    }
    __syncthreads();
    const double* U = Us + a * SV_LD * SV_LD;
    if (tid < nn) {
      const int b = tid % n, rr = tid / n;
      double s = 0;
      for (int q = 0; q < n; ++q) s += U[b * SV_LD + q] * Tp[q * SV_LD + rr];
      X[b * SV_LD + rr] = s;  // (U T')[b][r']
    }
    __syncthreads();
    double tval = 0;
    if (tid < nn) {
      const int i = tid % n, j = tid / n;  // message layout M[i + n j]
      double sij = 0, sji = 0;
      for (int q = 0; q < n; ++q) {
        sij += X[i * SV_LD + q] * U[j * SV_LD + q];
        sji += X[j * SV_LD + q] * U[i * SV_LD + q];
      }
      tval = 0.5 * (sij + sji);
    }
    const double n2 = block_sum(tval * tval, red);
    const double nrm = sqrt(n2);
    const double sc = (normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
    const long long off = tk->out[a];
    double dot = 0, o2 = 0, d2 = 0, bad = 0;
    if (tid < nn) {
      const double nv = tval * sc;
      const double ov = (double)msg_in[off + tid];
      if (!isfinite(nv)) bad = 1;
      dot = ov * nv;
      o2 = ov * ov;
      d2 = (nv - ov) * (nv - ov);
      msg_out[off + tid] = (TM)nv;
    }
    dot = block_sum(dot, red);
    o2 = block_sum(o2, red);
    d2 = block_sum(d2, red);
    bad = block_sum(bad, red);
    const int slot = tk->slot[a];
    if (tid == 0 && slot >= 0) {
      const double nn2 = n2 * sc * sc;
      MsgStat stt;
      stt.diff = 1.0 - (dot * dot) / (o2 * nn2);
      stt.res2 = nn2 > 0 ? d2 / nn2 : 0.0;
      stt.norm = nrm;
      stt.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
      stats[slot] = stt;
    }
    __syncthreads();
  }
}

// FP32 flavour of k_vertex_big on the matrix cores (v_mfma_f32_16x16x4f32).
// Step 3, the rotation of bond a, is per tile of 16 fibres the product
//   Y'[f, r] = sum_j Y[f, j] U[j, r]       (A: 16 fibres x 4 of j, B: U, K = n <= 16)
// and step 4, the Gram of bond a, is per group of 4 fibres
//   T'[r, r'] += sum_f (w_f Y[f, r]) Y[f, r']   (K = 4 fibres; one load per lane
//   feeds both operands, since A[r][k] and B[k][r'] hold the same fibre column).
// Each wave takes its own tiles (NT in flight per iteration, their loads issued
// together); the Gram accumulates in FP32 for 16 MFMAs (64 fibres) at a time and
// is carried in FP64 between them, the four waves' partial Grams summed in FP64.
// Lane layout (16x16x4 f32): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k,
// D[4 (l / 16) + v][l % 16] in lane l, value v.
#define F32_DROW(l, v) (4 * ((l) >> 4) + (v))
#define F32_DCOL(l, v) ((l) & 15)
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int VBM_CHUNK = 1024;  // fibres whose base offset and weight are tabulated at once
// fixed part + Jacobi work / wave partials + the chunk tables (base offsets, weights)
constexpr size_t VBM_LDS = VB_FIXED + 4 * SV_LD * SV_LD * 8 + (size_t)VBM_CHUNK * 8;
constexpr int VBM_TILES[3] = {1, 2, 4};                         // MFMA tiles in flight per wave
constexpr int VBM_FT = 2;
static_assert(4 * SV_LD * SV_LD * 8 >= 4 * 256 * 8, "the wave partial Grams must fit the Jacobi work area");

template <typename TS, typename TM, int NT>
__global__ __launch_bounds__(SV_THREADS) void k_vertex_big_mfma(const VDesc* __restrict__ vdesc,
                                                               const SmallTask* __restrict__ tasks,
                                                               const TS* __restrict__ psi, const TM* msg_in,
                                                               TM* msg_out, MsgStat* __restrict__ stats,
                                                               int normalize, float* __restrict__ scratch,
                                                               double* __restrict__ seeds, int warm,
                                                               const int* __restrict__ active,
                                                               const long long* __restrict__ sweep_ctr) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* Us = reinterpret_cast<double*>(smem);
  double* lam = Us + ITNBP_MAXK * SV_LD * SV_LD;
  double* rc = lam + ITNBP_MAXK * SV_MAXCHI;
  double* rs = rc + 32;
  int* rp = reinterpret_cast<int*>(rs + 32);
  int* rq = rp + 32;
  double* gpart = reinterpret_cast<double*>(rq + 32);
  double* Tp = gpart + SV_THREADS;
  double* X = Tp + SV_LD * SV_LD;
  double* red = X + SV_LD * SV_LD;
  double* As = reinterpret_cast<double*>(smem + VB_FIXED);  // Jacobi work (step 2), wave partials (step 4)
  int* cbase = reinterpret_cast<int*>(smem + VB_FIXED + 4 * SV_LD * SV_LD * 8);  // [VBM_CHUNK]
  float* cw = reinterpret_cast<float*>(cbase + VBM_CHUNK);                         // [VBM_CHUNK]

  __shared__ int dims[ITNBP_MAXR], strides[ITNBP_MAXR];
  const SmallTask* tk = tasks + blockIdx.x;
  if (active && !active[tk->inst]) return;
  if (sweep_ctr) warm = (*sweep_ctr % 16) != 0;
  const VDesc* V = vdesc + tk->v;
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63;
  const int li = l & 15, lk = l >> 4;
  const int k = V->k, D = V->D;
  if (tid == 0) {
    int s = 1;
    for (int a = 0; a <= k; ++a) {
      dims[a] = V->dims[a];
      strides[a] = s;
      s *= dims[a];
    }
  }
  __syncthreads();
  float* Y = scratch + tk->yoff;
  const TS* src = psi + V->psi_off;
  // 2. eigendecompositions, one bond per wave
  for (int a = wv; a < k; a += 4)
    bond_eigh(msg_in + V->in_msg[a + 1], dims[a + 1], As + wv * SV_LD * SV_LD, Us + a * SV_LD * SV_LD,
              lam + a * SV_MAXCHI, seeds + (long long)tk->useed[a] * SEED_LD * SEED_LD, warm, rc + 8 * wv,
              rs + 8 * wv, rp + 8 * wv, rq + 8 * wv);
  __threadfence_block();
  __syncthreads();
  // 3. rotate every bond in place (bond 0 reads psi), 16 fibres per MFMA tile
  for (int a = 0; a < k; ++a) {
    const int n = dims[a + 1], st = strides[a + 1];
    const int F = D / n;
    const double* U = Us + a * SV_LD * SV_LD;
    float bu[4];  // B operand: U[j][r], j = 4 kk + lk, r = li
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int j = 4 * kk + lk;
      bu[kk] = (j < n && li < n) ? (float)U[j * SV_LD + li] : 0.f;
    }
    const int ntiles = (F + 15) / 16;
    for (int t0 = wv * NT; t0 < ntiles; t0 += 4 * NT) {
      float av[NT][4];
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        const int f = (t0 + q) * 16 + li;
        const int base = f < F ? (f % st) + (f / st) * st * n : -1;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const int j = 4 * kk + lk;
          av[q][kk] = (base >= 0 && j < n) ? (a == 0 ? load_as<float>(src, base + j * st) : Y[base + j * st]) : 0.f;
        }
      }
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][kk], bu[kk], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int f = (t0 + q) * 16 + F32_DROW(l, v), r = F32_DCOL(l, v);
          if (f < F && r < n) Y[(f % st) + (f / st) * st * n + r * st] = acc[v];
        }
      }
    }
    __threadfence_block();
    __syncthreads();
  }
  // 4-5. per bond: weighted Gram, rotate back, normalise, statistics
  for (int a = 0; a < k; ++a) {
    const int n = dims[a + 1], st = strides[a + 1], nn = n * n;
    const int F = D / n;
    double g64[4] = {0, 0, 0, 0};
    // fibres in chunks: their base offsets and weights tabulated once per fibre
    // (not per lane), then wave wv takes the chunk's fibre groups (4 fibres per
    // MFMA) wv, wv + 4, ...; NT groups per iteration, in runs of 16 MFMAs
    // between FP64 flushes
    for (int c0 = 0; c0 < F; c0 += VBM_CHUNK) {
      const int nc = min(VBM_CHUNK, F - c0);
      for (int i = tid; i < nc; i += SV_THREADS) {
        const int f = c0 + i;
        const int base = (f % st) + (f / st) * st * n;
        double w = 1.0;
        for (int b = 0; b < k; ++b)
          if (b != a) w *= lam[b * SV_MAXCHI + (base / strides[b + 1]) % dims[b + 1]];
        cbase[i] = base;
        cw[i] = (float)w;
      }
      __syncthreads();
      const int ngroups = (nc + 3) / 4;
      for (int g0 = wv * NT; g0 < ngroups; g0 += 64 * NT) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int gg = g0; gg < min(ngroups, g0 + 64 * NT); gg += 4 * NT) {
          float y[NT], wy[NT];
#pragma unroll
          for (int q = 0; q < NT; ++q) {
            const int fl = (gg + q) * 4 + lk;
            y[q] = 0.f;
            wy[q] = 0.f;
            if (gg + q < ngroups && fl < nc && li < n) {
              y[q] = Y[cbase[fl] + li * st];
              wy[q] = cw[fl] * y[q];
            }
          }
#pragma unroll
          for (int q = 0; q < NT; ++q) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(wy[q], y[q], acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) g64[v] += (double)acc[v];
      }
      __syncthreads();
    }
    // the four waves' partial Grams in LDS (the Jacobi work area), then T'
    double* red4 = As;
#pragma unroll
    for (int v = 0; v < 4; ++v) red4[wv * 256 + F32_DROW(l, v) * 16 + F32_DCOL(l, v)] = g64[v];
    __syncthreads();
    if (tid < nn) {
      const int rr = tid % n, r2 = tid / n;
      Tp[rr * SV_LD + r2] = (red4[rr * 16 + r2] + red4[256 + rr * 16 + r2]) +
                            (red4[512 + rr * 16 + r2] + red4[768 + rr * 16 + r2]);  // T'[r][r']
    }
    __syncthreads();
    const double* U = Us + a * SV_LD * SV_LD;
    if (tid < nn) {
      const int b = tid % n, rr = tid / n;
      double s = 0;
      for (int q = 0; q < n; ++q) s += U[b * SV_LD + q] * Tp[q * SV_LD + rr];
      X[b * SV_LD + rr] = s;  // (U T')[b][r']
    }
    __syncthreads();
    double tval = 0;
    if (tid < nn) {
      const int i = tid % n, j = tid / n;  // message layout M[i + n j]
      double sij = 0, sji = 0;
      for (int q = 0; q < n; ++q) {
        sij += X[i * SV_LD + q] * U[j * SV_LD + q];
        sji += X[j * SV_LD + q] * U[i * SV_LD + q];
      }
      tval = 0.5 * (sij + sji);
    }
    const double n2 = block_sum(tval * tval, red);
    const double nrm = sqrt(n2);
    const double sc = (normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
    const long long off = tk->out[a];
    double dot = 0, o2 = 0, d2 = 0, bad = 0;
    if (tid < nn) {
      const double nv = tval * sc;
      const double ov = (double)msg_in[off + tid];
      if (!isfinite(nv)) bad = 1;
      dot = ov * nv;
      o2 = ov * ov;
      d2 = (nv - ov) * (nv - ov);
      msg_out[off + tid] = (TM)nv;
    }
    dot = block_sum(dot, red);
    o2 = block_sum(o2, red);
    d2 = block_sum(d2, red);
    bad = block_sum(bad, red);
    const int slot = tk->slot[a];
    if (tid == 0 && slot >= 0) {
      const double nn2 = n2 * sc * sc;
      MsgStat stt;
      stt.diff = 1.0 - (dot * dot) / (o2 * nn2);
      stt.res2 = nn2 > 0 ? d2 / nn2 : 0.0;
      stt.norm = nrm;
      stt.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
      stats[slot] = stt;
    }
    __syncthreads();
  }
}

struct SmallState {
  std::vector<SmallTask> tasks;
  DevBuf d_tasks;
  size_t lds = 0;
  int fmax = 0;  // largest fibre count over the tasks' bonds
  // sites too large for LDS: k_vertex_big with scratch tensors
  std::vector<SmallTask> big;
  DevBuf d_big, ybuf;
  long long ytotal = 0;  // scratch elements
  // eigenvector seeds, one SEED_LD x SEED_LD slot per message id; valid after the
  // first sweep since the last prepare, and every 16th sweep starts cold so that
  // rounding in the accumulated rotations cannot build up
  DevBuf seeds;
  long long sweeps = 0;
  // batched update: instance of every vertex (empty: 0), the per-instance
  // active flags the kernels skip finished instances by, and the device sweep
  // counter the warm start is taken from (both null outside the batched update)
  std::vector<int> vinst;
  const int* active = nullptr;
  const long long* sweep_ctr = nullptr;
  // k_vertex_big variant the autotuner chose (0: none yet) and for which flavour
  int ft = 0;
  bool ft_f32 = false;
  // persistent small-site launch (k_vertex_eig_queue): its task counter, its
  // grid of resident workgroups (0: not yet sized), and the autotuner's choice
  // between it and one workgroup per task (-1: untuned, one per task)
  DevBuf queue;
  int queue_grid = 0;
  int eig_queue = -1;
};

// ITNBP_VERTEX_QUEUE=0/1 forces the small-site launch form, else the
// autotuner's choice, else one workgroup per task
bool use_queue(const SmallState* st) {
  if (const char* e = getenv("ITNBP_VERTEX_QUEUE")) return atoi(e) != 0;
  return st->eig_queue == 1;
}

// the k_vertex_big variant of a launch: ITNBP_VERTEX_BIG_TILE=<tile> forces
// one (VB_TILES, or VBM_TILES for FP32), else the autotuner's choice for the
// flavour, else its default
template <typename TA>
int big_variant(const SmallState* st) {
  if (const char* e = getenv("ITNBP_VERTEX_BIG_TILE")) {
    const int f = atoi(e);
    for (int t : VB_TILES)
      if (f == t) return f;
    if (sizeof(TA) == 4)
      for (int t : VBM_TILES)
        if (f == t) return f;
  }
  if (st->ft > 0 && st->ft_f32 == (sizeof(TA) == 4)) return st->ft;
  return sizeof(TA) == 4 ? VBM_FT : VB_FT;
}

SmallState* small_of(itnbp_net* net) { return reinterpret_cast<SmallState*>(net->small); }

template <typename TS, typename TA, typename TM, int FT>
hipError_t launch_big(itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, void* msg_out,
                      MsgStat* stats, int normalize, double* seeds, int warm, hipStream_t stream) {
  auto kern = k_vertex_big<TS, TA, TM, FT>;
  const size_t lds = vb_lds<TA>(FT);
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((unsigned)st->big.size()), dim3(SV_THREADS), lds, stream, (const VDesc*)net->d_vdesc,
                     (const SmallTask*)st->d_big.p, (const TS*)psi, (const TM*)msg_in, (TM*)msg_out, stats, normalize,
                     (TA*)st->ybuf.p, seeds, warm, st->active, st->sweep_ctr);
  return hipGetLastError();
}

template <typename TS, typename TM, int NT>
hipError_t launch_big_mfma(itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, void* msg_out,
                           MsgStat* stats, int normalize, double* seeds, int warm, hipStream_t stream) {
  auto kern = k_vertex_big_mfma<TS, TM, NT>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)VBM_LDS);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((unsigned)st->big.size()), dim3(SV_THREADS), VBM_LDS, stream,
                     (const VDesc*)net->d_vdesc, (const SmallTask*)st->d_big.p, (const TS*)psi, (const TM*)msg_in,
                     (TM*)msg_out, stats, normalize, (float*)st->ybuf.p, seeds, warm, st->active, st->sweep_ctr);
  return hipGetLastError();
}

// ft: the VALU kernel's Gram tile (VB_TILES), or for the FP32 flavour the MFMA
// kernel with ft tiles in flight per wave (VBM_TILES)
template <typename TS, typename TA, typename TM>
hipError_t launch_big_ft(int ft, itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, void* msg_out,
                         MsgStat* stats, int normalize, double* seeds, int warm, hipStream_t stream) {
  if constexpr (sizeof(TA) == 4) {
    switch (ft) {
      case 1: return launch_big_mfma<TS, TM, 1>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm, stream);
      case 2: return launch_big_mfma<TS, TM, 2>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm, stream);
      case 4: return launch_big_mfma<TS, TM, 4>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm, stream);
      default: break;
    }
  }
  switch (ft) {
    case 160: return launch_big<TS, TA, TM, 160>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm, stream);
    case 256: return launch_big<TS, TA, TM, 256>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm, stream);
    default: return launch_big<TS, TA, TM, 192>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm, stream);
  }
}

// the LDS-resident sites: one workgroup per task, or the persistent queue form
template <typename TS, typename TA, typename TM>
hipError_t launch_small(itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, void* msg_out,
                        MsgStat* stats, int normalize, double* seeds, int warm, bool queue, hipStream_t stream) {
  const int ntasks = (int)st->tasks.size();
  if (!queue || !st->queue.p) {
    auto kern = k_vertex_eig<TS, TA, TM>;
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)st->lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3((unsigned)ntasks), dim3(SV_THREADS), st->lds, stream, (const VDesc*)net->d_vdesc,
                       (const SmallTask*)st->d_tasks.p, (const TS*)psi, (const TM*)msg_in, (TM*)msg_out, stats,
                       normalize, st->fmax, seeds, warm, st->active, st->sweep_ctr);
    return hipGetLastError();
  }
  auto kern = k_vertex_eig_queue<TS, TA, TM>;
  hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)st->lds);
  if (e != hipSuccess) return e;
  if (st->queue_grid == 0) {
    // every workgroup of the grid resident at once
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kern, SV_THREADS, st->lds);
    if (e != hipSuccess) return e;
    st->queue_grid = std::max(1, std::min(ntasks, std::max(1, per_cu) * net->ctx->num_cu));
  }
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, stream, (int*)st->queue.p);
  hipLaunchKernelGGL(kern, dim3((unsigned)st->queue_grid), dim3(SV_THREADS), st->lds, stream,
                     (const VDesc*)net->d_vdesc, (const SmallTask*)st->d_tasks.p, ntasks, (int*)st->queue.p,
                     (const TS*)psi, (const TM*)msg_in, (TM*)msg_out, stats, normalize, st->fmax, seeds, warm,
                     st->active, st->sweep_ctr);
  return hipGetLastError();
}

template <typename TS, typename TA, typename TM>
hipError_t launch_impl(itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, void* msg_out,
                       MsgStat* stats, int normalize) {
  const int warm = (st->sweeps % 16) != 0 ? 1 : 0;
  double* seeds = (double*)st->seeds.p;
  // the big sites (the longest workgroups) first, so that the LDS-resident
  // sites fill in behind them
  if (!st->big.empty()) {
    hipError_t e = launch_big_ft<TS, TA, TM>(big_variant<TA>(st), net, st, psi, msg_in, msg_out, stats, normalize,
                                            seeds, warm, net->ctx->stream);
    if (e != hipSuccess) return e;
  }
  if (!st->tasks.empty()) {
    hipError_t e = launch_small<TS, TA, TM>(net, st, psi, msg_in, msg_out, stats, normalize, seeds, warm,
                                            use_queue(st), net->ctx->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Times one launch of k_vertex_big with every VB_TILES tile (and, for FP32,
// the MFMA kernel with every VBM_TILES count of tiles in flight) on scratch
// outputs (messages, statistics and eigenvector seeds; the network's are
// untouched) and keeps the fastest in st->ft.
template <typename TS, typename TA, typename TM>
int autotune_big(itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, int normalize) {
  itnbp_ctx* ctx = net->ctx;
  DevBuf out, stats, seeds;
  auto cleanup = [&]() {
    free_buf(out);
    free_buf(stats);
    free_buf(seeds);
  };
  int rc = ensure_buf(ctx, out, (size_t)net->msg_total * sizeof(TM));
  if (rc == ITNBP_OK) rc = ensure_buf(ctx, stats, sizeof(MsgStat) * (size_t)std::max(1, 2 * net->ne));
  if (rc == ITNBP_OK) rc = ensure_buf(ctx, seeds, st->seeds.bytes);
  if (rc != ITNBP_OK) {
    cleanup();
    return rc;
  }
  float best = INFINITY;
  int best_ft = big_variant<TA>(st);
  std::vector<int> cand(VB_TILES, VB_TILES + 3);
  if (sizeof(TA) == 4) cand.insert(cand.end(), VBM_TILES, VBM_TILES + 3);
  for (int rep = 0; rep < 2; ++rep)  // the first round warms the code and caches
    for (int ft : cand) {
      hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
      if (e == hipSuccess)
        e = launch_big_ft<TS, TA, TM>(ft, net, st, psi, msg_in, out.p, (MsgStat*)stats.p, normalize,
                                      (double*)seeds.p, 0, ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
      if (e != hipSuccess) {
        cleanup();
        return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("vertex autotune: ") + hipGetErrorString(e));
      }
      if (rep == 1 && ms < best) {
        best = ms;
        best_ft = ft;
      }
    }
  cleanup();
  st->ft = best_ft;
  st->ft_f32 = sizeof(TA) == 4;
  return ITNBP_OK;
}

}  // namespace

bool small_path_enabled(const itnbp_net* net) {
  const char* e = getenv("ITNBP_DISABLE_VERTEX");
  return net->msgs_sym && !(e && e[0] == '1');
}

int small_prepare(itnbp_net* net, const std::vector<int>& seq, int acc_f64, std::vector<char>& handled) {
  itnbp_ctx* ctx = net->ctx;
  SmallState* st = small_of(net);
  if (!st) {
    st = new SmallState();
    net->small = st;
  }
  st->tasks.clear();
  st->big.clear();
  st->ytotal = 0;
  st->lds = 0;
  st->fmax = 0;
  st->sweeps = 0;
  if (!small_path_enabled(net)) return ITNBP_OK;
  std::vector<int> pos(2 * net->ne, -1);
  for (size_t i = 0; i < seq.size(); ++i) pos[seq[i]] = (int)i;
  const size_t eb = acc_f64 ? 8 : 4;
  int dmax_all = 0;
  for (int v = 0; v < net->nv; ++v) {
    const VDesc& V = net->vdesc[v];
    if (V.k < 1 || V.k > ITNBP_MAXK) continue;
    bool ok = true;
    int nmin = SV_MAXCHI;
    for (int a = 1; a <= V.k && ok; ++a) {
      ok = V.dims[a] <= SV_MAXCHI;
      nmin = std::min(nmin, V.dims[a]);
    }
    if (!ok) continue;
    // one LDS carve-up serves every task of the launch: the largest fibre table
    // and the largest site tensor
    const int fmax = std::max(st->fmax, V.D / std::max(1, nmin));
    const int dmax = std::max(dmax_all, V.D);
    const size_t lds = SV_FIXED + (((size_t)fmax * 12 + 15) & ~(size_t)15) + (size_t)dmax * eb;
    const bool fits = lds + 256 <= 160 * 1024;  // + the static LDS of the kernel
    SmallTask t{};
    t.v = v;
    t.inst = st->vinst.empty() ? 0 : st->vinst[v];
    for (int a = 1; a <= V.k && ok; ++a) {
      const int e = net->inc_edge[net->inc_ptr[v] + a - 1];
      const int w = net->edges[2 * e] == v ? net->edges[2 * e + 1] : net->edges[2 * e];
      const int m = msg_id_into(net, e, w);
      if (pos[m] < 0 || handled[pos[m]]) ok = false;
      else {
        t.out[a - 1] = net->msg_off[m];
        t.slot[a - 1] = pos[m];
        t.useed[a - 1] = msg_id_into(net, e, v);
      }
    }
    if (!ok) continue;
    for (int a = 0; a < V.k; ++a) handled[t.slot[a]] = 1;
    if (fits) {
      st->tasks.push_back(t);
      st->fmax = fmax;
      dmax_all = dmax;
      st->lds = lds;
    } else {
      t.yoff = st->ytotal;
      st->ytotal += V.D;
      st->big.push_back(t);
    }
  }
  // costliest sites first (rotations and Grams both scale as D x sum of the
  // bond dimensions), so that the hardware dispatcher starts the long
  // workgroups early and the short ones fill in behind them
  auto cost = [&](const SmallTask& t) {
    const VDesc& V = net->vdesc[t.v];
    long long s = 0;
    for (int a = 1; a <= V.k; ++a) s += V.dims[a];
    return (long long)V.D * s;
  };
  auto by_cost = [&](const SmallTask& x, const SmallTask& y) { return cost(x) > cost(y); };
  std::stable_sort(st->tasks.begin(), st->tasks.end(), by_cost);
  std::stable_sort(st->big.begin(), st->big.end(), by_cost);
  if (!st->tasks.empty() || !st->big.empty())
    ITNBP_RETCHK(ensure_buf(ctx, st->seeds, (size_t)std::max(1, 2 * net->ne) * SEED_LD * SEED_LD * 8));
  if (!st->big.empty()) {
    ITNBP_RETCHK(ensure_buf(ctx, st->d_big, st->big.size() * sizeof(SmallTask)));
    ITNBP_RETCHK(ensure_buf(ctx, st->ybuf, (size_t)st->ytotal * eb));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_big.p, st->big.data(), st->big.size() * sizeof(SmallTask),
                                     hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (!st->tasks.empty()) {
    ITNBP_RETCHK(ensure_buf(ctx, st->queue, 256));
    st->queue_grid = 0;
    ITNBP_RETCHK(ensure_buf(ctx, st->d_tasks, st->tasks.size() * sizeof(SmallTask)));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(st->d_tasks.p, st->tasks.data(), st->tasks.size() * sizeof(SmallTask),
                                     hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return ITNBP_OK;
}

void small_set_batch(itnbp_net* net, const std::vector<int>& vinst, const int* active, const long long* sweep_ctr) {
  SmallState* st = small_of(net);
  if (!st) {
    st = new SmallState();
    net->small = st;
  }
  st->vinst = vinst;
  st->active = active;
  st->sweep_ctr = sweep_ctr;
}

// Times one launch of the LDS-resident sites in each form (one workgroup per
// task; the persistent queue) on scratch outputs and keeps the faster.
template <typename TS, typename TA, typename TM>
int autotune_small(itnbp_net* net, SmallState* st, const void* psi, const void* msg_in, int normalize) {
  itnbp_ctx* ctx = net->ctx;
  DevBuf out, stats, seeds;
  auto cleanup = [&]() {
    free_buf(out);
    free_buf(stats);
    free_buf(seeds);
  };
  int rc = ensure_buf(ctx, out, (size_t)net->msg_total * sizeof(TM));
  if (rc == ITNBP_OK) rc = ensure_buf(ctx, stats, sizeof(MsgStat) * (size_t)std::max(1, 2 * net->ne));
  if (rc == ITNBP_OK) rc = ensure_buf(ctx, seeds, st->seeds.bytes);
  if (rc != ITNBP_OK) {
    cleanup();
    return rc;
  }
  float best = INFINITY;
  int best_q = 0;
  for (int rep = 0; rep < 2; ++rep)
    for (int q = 0; q < 2; ++q) {
      hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
      if (e == hipSuccess)
        e = launch_small<TS, TA, TM>(net, st, psi, msg_in, out.p, (MsgStat*)stats.p, normalize, (double*)seeds.p, 0,
                                     q == 1, ctx->stream);
      if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
      if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
      float ms = 0;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
      if (e != hipSuccess) {
        cleanup();
        return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("vertex autotune: ") + hipGetErrorString(e));
      }
      if (rep == 1 && ms < best) {
        best = ms;
        best_q = q;
      }
    }
  cleanup();
  st->eig_queue = best_q;
  return ITNBP_OK;
}

int small_autotune(itnbp_net* net, int acc_f64, const void* psi, int psi_dtype, const void* msg_in, int normalize) {
  SmallState* st = small_of(net);
  if (st && !st->tasks.empty() && st->queue.p) {
    int rc;
    if (acc_f64) {
      switch (psi_dtype) {
        case ITNBP_F64: rc = autotune_small<double, double, double>(net, st, psi, msg_in, normalize); break;
        case ITNBP_F32: rc = autotune_small<float, double, double>(net, st, psi, msg_in, normalize); break;
        default: rc = autotune_small<bf16_t, double, double>(net, st, psi, msg_in, normalize); break;
      }
    } else {
      switch (psi_dtype) {
        case ITNBP_F64: rc = autotune_small<double, float, float>(net, st, psi, msg_in, normalize); break;
        case ITNBP_F32: rc = autotune_small<float, float, float>(net, st, psi, msg_in, normalize); break;
        default: rc = autotune_small<bf16_t, float, float>(net, st, psi, msg_in, normalize); break;
      }
    }
    if (rc != ITNBP_OK) return rc;
  }
  if (!st || st->big.empty()) return ITNBP_OK;
  if (acc_f64) {
    switch (psi_dtype) {
      case ITNBP_F64: return autotune_big<double, double, double>(net, st, psi, msg_in, normalize);
      case ITNBP_F32: return autotune_big<float, double, double>(net, st, psi, msg_in, normalize);
      default: return autotune_big<bf16_t, double, double>(net, st, psi, msg_in, normalize);
    }
  }
  switch (psi_dtype) {
    case ITNBP_F64: return autotune_big<double, float, float>(net, st, psi, msg_in, normalize);
    case ITNBP_F32: return autotune_big<float, float, float>(net, st, psi, msg_in, normalize);
    default: return autotune_big<bf16_t, float, float>(net, st, psi, msg_in, normalize);
  }
}

int small_big_tile(itnbp_net* net) {
  SmallState* st = small_of(net);
  return st && !st->big.empty() ? st->ft : 0;
}

bool small_queue_used(itnbp_net* net) {
  SmallState* st = small_of(net);
  return st && !st->tasks.empty() && st->queue.p && use_queue(st);
}

const char* small_big_kind(itnbp_net* net) {
  SmallState* st = small_of(net);
  if (!st || st->big.empty() || st->ft == 0) return "";
  for (int t : VBM_TILES)
    if (st->ft_f32 && st->ft == t) return "mfma";
  return "valu";
}

int small_task_count(itnbp_net* net) {
  SmallState* st = small_of(net);
  return st ? (int)(st->tasks.size() + st->big.size()) : 0;
}

int small_sweep(itnbp_net* net, int acc_f64, const void* psi, int psi_dtype, const void* msg_in, void* msg_out,
                MsgStat* stats, int normalize) {
  itnbp_ctx* ctx = net->ctx;
  SmallState* st = small_of(net);
  if (!st || (st->tasks.empty() && st->big.empty())) return ITNBP_OK;
  const int e0 = prof_mark(ctx);
  hipError_t e;
  if (acc_f64) {
    switch (psi_dtype) {
      case ITNBP_F64: e = launch_impl<double, double, double>(net, st, psi, msg_in, msg_out, stats, normalize); break;
      case ITNBP_F32: e = launch_impl<float, double, double>(net, st, psi, msg_in, msg_out, stats, normalize); break;
      default: e = launch_impl<bf16_t, double, double>(net, st, psi, msg_in, msg_out, stats, normalize); break;
    }
  } else {
    switch (psi_dtype) {
      case ITNBP_F64: e = launch_impl<double, float, float>(net, st, psi, msg_in, msg_out, stats, normalize); break;
      case ITNBP_F32: e = launch_impl<float, float, float>(net, st, psi, msg_in, msg_out, stats, normalize); break;
      default: e = launch_impl<bf16_t, float, float>(net, st, psi, msg_in, msg_out, stats, normalize); break;
    }
  }
  prof_add(ctx, "vertex_eig", e0, prof_mark(ctx));
  ITNBP_HIPCHK(ctx, e);
  ++st->sweeps;
  return ITNBP_OK;
}

void release_small_state(itnbp_net* net) {
  SmallState* st = small_of(net);
  if (!st) return;
  free_buf(st->d_tasks);
  free_buf(st->queue);
  free_buf(st->d_big);
  free_buf(st->ybuf);
  free_buf(st->seeds);
  delete st;
  net->small = nullptr;
}
