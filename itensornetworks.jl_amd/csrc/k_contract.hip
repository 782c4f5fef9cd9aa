// Generic "open contraction" kernels: absorb incoming BP messages into a site
// tensor by mode products, contract with the conjugate site tensor over every
// axis but the open ones, normalise, and record message_diff / residual.
//
// Semantics: updated_message(Algorithm"contract", bpc, e)
// (src/caches/abstractbeliefpropagationcache.jl:225-239) on the <psi|psi>
// network, with the bra = conj(psi) generated on the fly and the identity
// operator elided; message_diff as in :32-36.
//
// Two paths:
//   small - one workgroup per task, psi read from HBM, the mode-product
//           intermediates ping-pong in LDS (used when 2*D + chi^2 + P^2
//           accumulator elements fit in 160 KiB);
//   large - mode products through a global scratch (one launch per absorbed
//           axis over every task of a batch), then a split-K contraction with
//           LDS tiles and a deterministic finaliser.
#include <cstdlib>

#include "itnbp_internal.hpp"

namespace {

struct OcGeom {
  int D, ds, Aj, nj, lo_n, Bj, P, R;
};

__host__ __device__ inline OcGeom oc_geom(const VDesc& V, const OcTask& t) {
  OcGeom g;
  g.D = V.D;
  g.ds = t.s_open ? V.dims[0] : 1;
  if (t.j > 0) {
    int A = 1;
    for (int a = 0; a < t.j; ++a) A *= V.dims[a];
    g.Aj = A;
    g.nj = V.dims[t.j];
  } else {
    g.Aj = V.D;
    g.nj = 1;
  }
  g.lo_n = g.Aj / g.ds;
  g.Bj = V.D / (g.Aj * g.nj);
  g.P = g.ds * g.nj;
  g.R = g.lo_n * g.Bj;
  return g;
}

__host__ __device__ inline int max_absorbed_dim(const VDesc& V, int mask) {
  int m = 1;
  for (int a = 1; a <= V.k; ++a)
    if ((mask >> a) & 1) m = V.dims[a] > m ? V.dims[a] : m;
  return m;
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// sum over the whole block (blockDim.x a multiple of 64); red holds >= 4 T
template <typename T>
__device__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

// Normalise T (P2 values in LDS), compare with the old message, write out.
template <typename TM, typename TA>
__device__ void finalize_task(const OcTask& t, const TA* Tb, int P2, const TM* msg_in, TM* out,
                              MsgStat* stats, double* redd) {
  // the squared norm in FP64 whatever TA: an FP32 square underflows below 1e-19
  double n2 = 0;
  for (int o = threadIdx.x; o < P2; o += blockDim.x) n2 += (double)Tb[o] * (double)Tb[o];
  n2 = block_sum(n2, redd);
  const double nrm = sqrt(n2);
  const double sc = (t.normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
  double dot = 0, o2 = 0, d2 = 0, bad = 0;
  for (int o = threadIdx.x; o < P2; o += blockDim.x) {
    const double nv = (double)Tb[o] * sc;
    if (!isfinite(nv)) bad = 1;
    if (t.old_off >= 0) {
      const double ov = (double)msg_in[t.old_off + o];
      dot += ov * nv;
      o2 += ov * ov;
      d2 += (nv - ov) * (nv - ov);
    }
  }
  // all threads have read the old message before anyone overwrites it
  __syncthreads();
  for (int o = threadIdx.x; o < P2; o += blockDim.x) out[t.out_off + o] = (TM)((double)Tb[o] * sc);
  dot = block_sum(dot, redd);
  o2 = block_sum(o2, redd);
  d2 = block_sum(d2, redd);
  bad = block_sum(bad, redd);
  if (threadIdx.x == 0 && t.slot >= 0) {
    const double nn2 = (double)n2 * sc * sc;
    MsgStat st;
    st.diff = (t.old_off >= 0) ? 1.0 - (dot * dot) / (o2 * nn2) : 0.0;
    st.res2 = (t.old_off >= 0) ? (nn2 > 0 ? d2 / nn2 : 0.0) : 0.0;
    st.norm = nrm;
    st.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
    stats[t.slot] = st;
  }
}

// ---------------------------------------------------------------------------
// small path
// ---------------------------------------------------------------------------
template <typename TP, typename TM, typename TA>
__global__ __launch_bounds__(ITNBP_BLOCK) void k_oc_small(const VDesc* __restrict__ vds,
                                                          const OcTask* __restrict__ tasks,
                                                          const TP* __restrict__ psi,
                                                          const TM* msg_in, TM* out,
                                                          MsgStat* __restrict__ stats) {
  extern __shared__ __align__(16) unsigned char smem[];
  const OcTask t = tasks[blockIdx.x];
  const VDesc& V = vds[t.v];  // by reference: a copy indexed at run time lands in scratch
  const OcGeom g = oc_geom(V, t);
  const int D = V.D;
  const int nmax = max_absorbed_dim(V, t.absorb_mask);
  double* red = reinterpret_cast<double*>(smem);  // 16 doubles
  // the two ping-pong buffers, selected by arithmetic rather than a run-time
  // indexed pointer array (which would live in scratch)
  TA* const buf0 = reinterpret_cast<TA*>(smem + 128);
  auto bufp = [&](int i) { return buf0 + (size_t)i * D; };
  TA* Mst = bufp(2);
  TA* Tb = Mst + nmax * nmax;
  const TP* ps = psi + V.psi_off;
  const int tid = threadIdx.x;

  int cur = -1;  // -1: the current tensor is psi in HBM
  int A = 1;
  for (int a = 1; a <= V.k; ++a) {
    A *= V.dims[a - 1];
    if (!((t.absorb_mask >> a) & 1)) continue;
    const int n = V.dims[a];
    const TM* M = msg_in + V.in_msg[a];
    for (int i = tid; i < n * n; i += blockDim.x) Mst[i] = (TA)M[i];
    __syncthreads();
    TA* Y = bufp((cur + 1) & 1);
    const int An = A * n;
    for (int i = tid; i < D; i += blockDim.x) {
      const int lo = i % A;
      const int rest = i / A;
      const int c = rest % n;
      const int hi = rest / n;
      const int base = lo + An * hi;
      const TA* Mc = Mst + n * c;
      TA acc = 0;
      if (cur < 0) {
        for (int x = 0; x < n; ++x) acc += load_as<TA>(ps, base + A * x) * Mc[x];
      } else {
        const TA* X = bufp(cur);
        for (int x = 0; x < n; ++x) acc += X[base + A * x] * Mc[x];
      }
      Y[i] = acc;
    }
    cur = (cur + 1) & 1;
    __syncthreads();
  }

  // T[row, col] = sum_r Phi[offx(row) + roff(r)] * psi[offx(col) + roff(r)]
  const int P2 = g.P * g.P;
  int G = 1;
  while (G * 2 * P2 <= (int)blockDim.x && G < 64) G *= 2;
  const int per = blockDim.x / G;
  const TA* Phi = cur < 0 ? nullptr : bufp(cur);
  const int strideHi = g.Aj * g.nj;
  for (int ob = 0; ob < P2; ob += per) {
    const int o = ob + tid / G;
    const int gi = tid % G;
    TA acc = 0;
    if (o < P2) {
      const int row = o % g.P, col = o / g.P;
      const int offx = (row % g.ds) + g.Aj * (row / g.ds);
      const int offy = (col % g.ds) + g.Aj * (col / g.ds);
      for (int r = gi; r < g.R; r += G) {
        const int lo = r % g.lo_n, hi = r / g.lo_n;
        const int roff = g.ds * lo + strideHi * hi;
        const TA ph = Phi ? Phi[offx + roff] : load_as<TA>(ps, offx + roff);
        acc += ph * load_as<TA>(ps, offy + roff);
      }
    }
    for (int m = G / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m);
    if (o < P2 && gi == 0) Tb[o] = acc;
  }
  __syncthreads();
  finalize_task<TM, TA>(t, Tb, P2, msg_in, out, stats, red);
}

// ---------------------------------------------------------------------------
// large path
// ---------------------------------------------------------------------------
__device__ inline int nth_absorbed_axis(const VDesc& V, int mask, int level) {
  int c = 0;
  for (int a = 1; a <= V.k; ++a)
    if ((mask >> a) & 1) {
      if (c == level) return a;
      ++c;
    }
  return 0;
}

__host__ __device__ inline int popcount_bonds(int mask) {
  int c = 0;
  for (int a = 1; a < 32; ++a) c += (mask >> a) & 1;
  return c;
}

// one mode product of every task that has a level-th absorbed axis
template <typename TP, typename TM, typename TA>
__global__ __launch_bounds__(ITNBP_BLOCK) void k_mp_level(const VDesc* __restrict__ vds,
                                                          const OcTask* __restrict__ tasks,
                                                          int level, const TP* __restrict__ psi,
                                                          const TM* __restrict__ msg_in,
                                                          TA* __restrict__ scratch,
                                                          long long stride) {
  __shared__ TA Mst[64 * 64];
  const OcTask t = tasks[blockIdx.y];
  const VDesc& V = vds[t.v];  // by reference: a copy indexed at run time lands in scratch
  const int a = nth_absorbed_axis(V, t.absorb_mask, level);
  if (a == 0) return;
  int A = 1;
  for (int b = 0; b < a; ++b) A *= V.dims[b];
  const int n = V.dims[a];
  const TM* M = msg_in + V.in_msg[a];
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) Mst[i] = (TA)M[i];
  __syncthreads();
  TA* S0 = scratch + (long long)blockIdx.y * 2 * stride;
  TA* Y = S0 + (level & 1) * stride;
  const TA* X = S0 + ((level - 1) & 1) * stride;
  const TP* ps = psi + V.psi_off;
  const int D = V.D;
  const int An = A * n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < D;
       i += (long long)gridDim.x * blockDim.x) {
    const int ii = (int)i;
    const int lo = ii % A;
    const int rest = ii / A;
    const int c = rest % n;
    const int hi = rest / n;
    const int base = lo + An * hi;
    const TA* Mc = Mst + n * c;
    TA acc = 0;
    if (level == 0) {
      for (int x = 0; x < n; ++x) acc += load_as<TA>(ps, base + A * x) * Mc[x];
    } else {
      for (int x = 0; x < n; ++x) acc += X[base + A * x] * Mc[x];
    }
    Y[ii] = acc;
  }
}

// LDS-tiled variant of k_mp_level: a tile is either TH consecutive `hi` slabs
// (A n contiguous elements each) or, for large A n, a run of TL consecutive `lo`
// values for every x at one `hi`; the tile is loaded coalesced into LDS and
// every output reads its n inputs from there.
#define MP_TILE 4096
template <typename TP, typename TM, typename TA>
__global__ __launch_bounds__(ITNBP_BLOCK) void k_mp_tiled(const VDesc* __restrict__ vds,
                                                          const OcTask* __restrict__ tasks,
                                                          int level, const TP* __restrict__ psi,
                                                          const TM* __restrict__ msg_in,
                                                          TA* __restrict__ scratch,
                                                          long long stride) {
  __shared__ TA Xt[MP_TILE];
  __shared__ TA Mst[64 * 64];
  const OcTask t = tasks[blockIdx.y];
  const VDesc& V = vds[t.v];  // by reference: a copy indexed at run time lands in scratch
  const int a = nth_absorbed_axis(V, t.absorb_mask, level);
  if (a == 0) return;
  int A = 1;
  for (int b = 0; b < a; ++b) A *= V.dims[b];
  const int n = V.dims[a];
  const int B = V.D / (A * n);
  const TM* M = msg_in + V.in_msg[a];
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) Mst[i] = (TA)M[i];
  TA* S0 = scratch + (long long)blockIdx.y * 2 * stride;
  TA* Y = S0 + (level & 1) * stride;
  const TA* Xs = S0 + ((level - 1) & 1) * stride;
  const TP* ps = psi + V.psi_off;
  const int An = A * n;
  if (An <= MP_TILE) {
    const int TH = MP_TILE / An;
    const int ntiles = (B + TH - 1) / TH;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
      const int hi0 = tile * TH;
      const int th = min(TH, B - hi0);
      const int len = th * An;
      const long long base = (long long)hi0 * An;
      __syncthreads();
      for (int i = threadIdx.x; i < len; i += blockDim.x)
        Xt[i] = level == 0 ? load_as<TA>(ps, base + i) : Xs[base + i];
      __syncthreads();
      for (int o = threadIdx.x; o < len; o += blockDim.x) {
        const int lo = o % A, rest = o / A, c = rest % n, hl = rest / n;
        const TA* xr = Xt + lo + An * hl;
        const TA* mc = Mst + n * c;
        TA acc = 0;
        for (int x = 0; x < n; ++x) acc += xr[A * x] * mc[x];
        Y[base + o] = acc;
      }
    }
  } else {
    const int TL = MP_TILE / n;
    const int ntl = (A + TL - 1) / TL;
    const int ntiles = ntl * B;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
      const int hi = tile / ntl;
      const int lo0 = (tile % ntl) * TL;
      const int tl = min(TL, A - lo0);
      __syncthreads();
      for (int i = threadIdx.x; i < n * tl; i += blockDim.x) {
        const int x = i / tl, j = i % tl;
        const long long src = lo0 + j + (long long)A * (x + (long long)n * hi);
        Xt[x * TL + j] = level == 0 ? load_as<TA>(ps, src) : Xs[src];
      }
      __syncthreads();
      for (int o = threadIdx.x; o < n * tl; o += blockDim.x) {
        const int c = o / tl, j = o % tl;
        const TA* mc = Mst + n * c;
        TA acc = 0;
        for (int x = 0; x < n; ++x) acc += Xt[x * TL + j] * mc[x];
        Y[lo0 + j + (long long)A * (c + (long long)n * hi)] = acc;
      }
    }
  }
}

#define OC_RC 32  // rest indices per LDS tile

// partial T over a slice of the rest index for every task
template <typename TP, typename TA>
__global__ __launch_bounds__(ITNBP_BLOCK) void k_oc_partial(const VDesc* __restrict__ vds,
                                                            const OcTask* __restrict__ tasks,
                                                            const TP* __restrict__ psi,
                                                            const TA* __restrict__ scratch,
                                                            long long stride,
                                                            TA* __restrict__ partials,
                                                            long long part_stride) {
  __shared__ TA tPhi[64 * OC_RC];
  __shared__ TA tPsi[64 * OC_RC];
  const OcTask t = tasks[blockIdx.y];
  const VDesc& V = vds[t.v];  // by reference: a copy indexed at run time lands in scratch
  const OcGeom g = oc_geom(V, t);
  const int nabs = popcount_bonds(t.absorb_mask);
  const TA* S0 = scratch + (long long)blockIdx.y * 2 * stride;
  const TA* Phi = nabs > 0 ? S0 + ((nabs - 1) & 1) * stride : nullptr;
  const TP* ps = psi + V.psi_off;
  const int P = g.P, P2 = P * P;
  const int nsplit = gridDim.x;
  const int per = (g.R + nsplit - 1) / nsplit;
  const int r0 = blockIdx.x * per;
  const int r1 = min(g.R, r0 + per);
  const int strideHi = g.Aj * g.nj;
  constexpr int MAXO = 16;  // outputs per thread (P2 <= 4096)
  TA acc[MAXO];
  for (int q = 0; q < MAXO; ++q) acc[q] = 0;
  for (int rb = r0; rb < r1; rb += OC_RC) {
    const int rc = min(OC_RC, r1 - rb);
    __syncthreads();
    for (int idx = threadIdx.x; idx < P * OC_RC; idx += blockDim.x) {
      const int rl = idx % OC_RC, x = idx / OC_RC;
      TA vphi = 0, vpsi = 0;
      if (rl < rc) {
        const int r = rb + rl;
        const int lo = r % g.lo_n, hi = r / g.lo_n;
        const int off = (x % g.ds) + g.Aj * (x / g.ds) + g.ds * lo + strideHi * hi;
        vpsi = load_as<TA>(ps, off);
        vphi = Phi ? Phi[off] : vpsi;
      }
      tPhi[x * OC_RC + rl] = vphi;
      tPsi[x * OC_RC + rl] = vpsi;
    }
    __syncthreads();
    for (int q = 0; q < MAXO; ++q) {
      const int o = threadIdx.x + q * blockDim.x;
      if (o >= P2) break;
      const int row = o % P, col = o / P;
      const TA* a = tPhi + row * OC_RC;
      const TA* b = tPsi + col * OC_RC;
      TA s = 0;
#pragma unroll 8
      for (int rl = 0; rl < OC_RC; ++rl) s += a[rl] * b[rl];
      acc[q] += s;
    }
  }
  TA* out = partials + (long long)blockIdx.y * part_stride + (long long)blockIdx.x * P2;
  for (int q = 0; q < MAXO; ++q) {
    const int o = threadIdx.x + q * blockDim.x;
    if (o >= P2) break;
    out[o] = acc[q];
  }
}

template <typename TM, typename TA>
__global__ __launch_bounds__(ITNBP_BLOCK) void k_oc_finalize(const VDesc* __restrict__ vds,
                                                             const OcTask* __restrict__ tasks,
                                                             const TA* __restrict__ partials,
                                                             long long part_stride, int nsplit,
                                                             const TM* msg_in, TM* out,
                                                             MsgStat* __restrict__ stats) {
  __shared__ TA Tb[64 * 64];
  __shared__ double red[8];
  const OcTask t = tasks[blockIdx.x];
  const VDesc& V = vds[t.v];  // by reference: a copy indexed at run time lands in scratch
  const OcGeom g = oc_geom(V, t);
  const int P2 = g.P * g.P;
  const TA* pp = partials + (long long)blockIdx.x * part_stride;
  for (int o = threadIdx.x; o < P2; o += blockDim.x) {
    TA s = 0;
    for (int q = 0; q < nsplit; ++q) s += pp[(long long)q * P2 + o];
    Tb[o] = s;
  }
  __syncthreads();
  finalize_task<TM, TA>(t, Tb, P2, msg_in, out, stats, red);
}

// ---------------------------------------------------------------------------
// small utilities
// ---------------------------------------------------------------------------
__global__ void k_reduce_stats(const MsgStat* __restrict__ st, int n, double* red) {
  __shared__ double sh[8];
  double s = 0, mres = 0, bad = 0, mdiff = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s += st[i].diff;
    mres = fmax(mres, st[i].res2);
    mdiff = fmax(mdiff, st[i].diff);
    bad = fmax(bad, st[i].bad);
    if (isnan(st[i].diff)) bad = fmax(bad, 0.5);
  }
  s = block_sum(s, sh);
  // max reductions
  for (int m = 32; m > 0; m >>= 1) {
    mres = fmax(mres, __shfl_xor(mres, m));
    bad = fmax(bad, __shfl_xor(bad, m));
    mdiff = fmax(mdiff, __shfl_xor(mdiff, m));
  }
  __shared__ double shm[3][4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    shm[0][w] = mres;
    shm[1][w] = bad;
    shm[2][w] = mdiff;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0, c = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      a = fmax(a, shm[0][i]);
      b = fmax(b, shm[1][i]);
      c = fmax(c, shm[2][i]);
    }
    red[0] = s;
    red[1] = a;
    red[2] = b;
    red[3] = c;
  }
}

__global__ void k_edge_scalars(const double* __restrict__ msg, const long long* __restrict__ offs,
                               const int* __restrict__ chis, int ne, double* out) {
  __shared__ double red[4];
  const int e = blockIdx.x;
  const int n2 = chis[e] * chis[e];
  const double* a = msg + offs[2 * e];
  const double* b = msg + offs[2 * e + 1];
  double s = 0;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) s += a[i] * b[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[e] = s;
}

__global__ void k_rdm2_combine(const double* __restrict__ half, const long long* __restrict__ au_off,
                               const long long* __restrict__ av_off, const int* __restrict__ dus,
                               const int* __restrict__ dvs, const int* __restrict__ chis, int n,
                               double* out) {
  __shared__ double red[4];
  const int p = blockIdx.x;
  const int du = dus[p], dv = dvs[p], chi = chis[p];
  const double* Au = half + au_off[p];
  const double* Av = half + av_off[p];
  const int lu = du * chi, lv = dv * chi;
  const int dd = du * dv;
  double* o = out + (long long)p * dd * dd;
  for (int q = 0; q < dd * dd; ++q) {
    const int row = q % dd, col = q / dd;
    const int su = row % du, sv = row / du, su2 = col % du, sv2 = col / du;
    double s = 0;
    for (int bb = threadIdx.x; bb < chi * chi; bb += blockDim.x) {
      const int b = bb % chi, b2 = bb / chi;
      s += Au[(su + du * b) + (long long)lu * (su2 + du * b2)] *
           Av[(sv + dv * b) + (long long)lv * (sv2 + dv * b2)];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) o[q] = s;
  }
}

__global__ void k_d2f(const double* __restrict__ s, float* __restrict__ d, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}
__global__ void k_f2d(const float* __restrict__ s, double* __restrict__ d, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    d[i] = (double)s[i];
}
template <typename TP>
__global__ void k_to_f64(const TP* __restrict__ s, double* __restrict__ d, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    d[i] = load_as<double>(s, i);
}
template <typename TP>
__global__ void k_scale_psi(TP* psi, const long long* __restrict__ offs, const int* __restrict__ sizes,
                            const double* __restrict__ scale) {
  const int v = blockIdx.y;
  TP* p = psi + offs[v];
  const double sc = scale[v];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < sizes[v]; i += gridDim.x * blockDim.x)
    p[i] = (TP)((double)p[i] * sc);
}
template <>
__global__ void k_scale_psi<bf16_t>(bf16_t* psi, const long long* __restrict__ offs,
                                    const int* __restrict__ sizes,
                                    const double* __restrict__ scale) {
  const int v = blockIdx.y;
  bf16_t* p = psi + offs[v];
  const double sc = scale[v];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < sizes[v]; i += gridDim.x * blockDim.x)
    p[i] = f32_to_bf16((float)((double)bf16_to_f32(p[i]) * sc));
}

template <typename TP, typename TM, typename TA>
hipError_t small_impl(const VDesc* d_vd, const OcTask* d_tasks, int n, const void* psi,
                      const void* msg_in, void* out, MsgStat* stats, size_t lds, hipStream_t s) {
  auto kern = k_oc_small<TP, TM, TA>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(n), dim3(ITNBP_BLOCK), lds, s, d_vd, d_tasks, (const TP*)psi,
                     (const TM*)msg_in, (TM*)out, stats);
  return hipGetLastError();
}

template <typename TP, typename TM, typename TA>
hipError_t large_impl(const VDesc* d_vd, const OcTask* d_tasks, const OcTask* h_tasks,
                      const VDesc* h_vd, int n, const void* psi, const void* msg_in, void* out,
                      MsgStat* stats, void* scratch, long long stride, void* partials, int nsplit,
                      long long part_stride, hipStream_t s) {
  int maxlev = 0;
  long long maxD = 0;
  for (int i = 0; i < n; ++i) {
    maxlev = std::max(maxlev, popcount_bonds(h_tasks[i].absorb_mask));
    maxD = std::max<long long>(maxD, h_vd[h_tasks[i].v].D);
  }
  const int chunks = (int)std::min<long long>(1024, (maxD + ITNBP_BLOCK * 4 - 1) / (ITNBP_BLOCK * 4));
  static const bool untiled = getenv("ITNBP_MP_UNTILED") != nullptr;
  for (int lev = 0; lev < maxlev; ++lev) {
    if (untiled)
      hipLaunchKernelGGL((k_mp_level<TP, TM, TA>), dim3(std::max(1, chunks), n), dim3(ITNBP_BLOCK), 0, s,
                         d_vd, d_tasks, lev, (const TP*)psi, (const TM*)msg_in, (TA*)scratch, stride);
    else
      hipLaunchKernelGGL((k_mp_tiled<TP, TM, TA>), dim3(std::max(1, std::min(256, (int)(maxD / MP_TILE) + 1)), n),
                         dim3(ITNBP_BLOCK), 0, s, d_vd, d_tasks, lev, (const TP*)psi, (const TM*)msg_in,
                         (TA*)scratch, stride);
  }
  hipLaunchKernelGGL((k_oc_partial<TP, TA>), dim3(nsplit, n), dim3(ITNBP_BLOCK), 0, s, d_vd, d_tasks,
                     (const TP*)psi, (const TA*)scratch, stride, (TA*)partials, part_stride);
  hipLaunchKernelGGL((k_oc_finalize<TM, TA>), dim3(n), dim3(ITNBP_BLOCK), 0, s, d_vd, d_tasks,
                     (const TA*)partials, part_stride, nsplit, (const TM*)msg_in, (TM*)out, stats);
  return hipGetLastError();
}

}  // namespace

size_t oc_small_lds_bytes(const VDesc& vd, const OcTask& t, int elem_bytes) {
  const OcGeom g = oc_geom(vd, t);
  const int nmax = max_absorbed_dim(vd, t.absorb_mask);
  const size_t elems = 2 * (size_t)vd.D + (size_t)nmax * nmax + (size_t)g.P * g.P;
  const size_t bytes = elems * (size_t)elem_bytes + 128;
  return bytes <= 160 * 1024 ? bytes : 0;
}

hipError_t launch_oc_small(int psi_dtype, int acc_f64, const VDesc* d_vd, const OcTask* d_tasks,
                           int n, const void* psi, const void* msg_in, void* out, MsgStat* stats,
                           size_t lds, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (acc_f64) {
    switch (psi_dtype) {
      case ITNBP_F64: return small_impl<double, double, double>(d_vd, d_tasks, n, psi, msg_in, out, stats, lds, s);
      case ITNBP_F32: return small_impl<float, double, double>(d_vd, d_tasks, n, psi, msg_in, out, stats, lds, s);
      default: return small_impl<bf16_t, double, double>(d_vd, d_tasks, n, psi, msg_in, out, stats, lds, s);
    }
  } else {
    switch (psi_dtype) {
      case ITNBP_F64: return small_impl<double, float, float>(d_vd, d_tasks, n, psi, msg_in, out, stats, lds, s);
      case ITNBP_F32: return small_impl<float, float, float>(d_vd, d_tasks, n, psi, msg_in, out, stats, lds, s);
      default: return small_impl<bf16_t, float, float>(d_vd, d_tasks, n, psi, msg_in, out, stats, lds, s);
    }
  }
}

hipError_t launch_oc_large(int psi_dtype, int acc_f64, const VDesc* d_vd, const OcTask* d_tasks,
                           const OcTask* h_tasks, const VDesc* h_vd, int n, const void* psi,
                           const void* msg_in, void* out, MsgStat* stats, void* scratch,
                           long long stride, void* partials, int nsplit, long long part_stride,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (acc_f64) {
    switch (psi_dtype) {
      case ITNBP_F64: return large_impl<double, double, double>(d_vd, d_tasks, h_tasks, h_vd, n, psi, msg_in, out, stats, scratch, stride, partials, nsplit, part_stride, s);
      case ITNBP_F32: return large_impl<float, double, double>(d_vd, d_tasks, h_tasks, h_vd, n, psi, msg_in, out, stats, scratch, stride, partials, nsplit, part_stride, s);
      default: return large_impl<bf16_t, double, double>(d_vd, d_tasks, h_tasks, h_vd, n, psi, msg_in, out, stats, scratch, stride, partials, nsplit, part_stride, s);
    }
  } else {
    switch (psi_dtype) {
      case ITNBP_F64: return large_impl<double, float, float>(d_vd, d_tasks, h_tasks, h_vd, n, psi, msg_in, out, stats, scratch, stride, partials, nsplit, part_stride, s);
      case ITNBP_F32: return large_impl<float, float, float>(d_vd, d_tasks, h_tasks, h_vd, n, psi, msg_in, out, stats, scratch, stride, partials, nsplit, part_stride, s);
      default: return large_impl<bf16_t, float, float>(d_vd, d_tasks, h_tasks, h_vd, n, psi, msg_in, out, stats, scratch, stride, partials, nsplit, part_stride, s);
    }
  }
}

hipError_t launch_reduce_stats(const MsgStat* stats, int n, double* red, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_stats, dim3(1), dim3(ITNBP_BLOCK), 0, s, stats, n, red);
  return hipGetLastError();
}

hipError_t launch_edge_scalars(const double* msg, const long long* offs, const int* chis, int ne,
                               double* out, hipStream_t s) {
  if (ne <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_edge_scalars, dim3(ne), dim3(ITNBP_BLOCK), 0, s, msg, offs, chis, ne, out);
  return hipGetLastError();
}

hipError_t launch_rdm2_combine(const double* half, const long long* au_off, const long long* av_off,
                               const int* du, const int* dv, const int* chis, int n, double* out,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rdm2_combine, dim3(n), dim3(ITNBP_BLOCK), 0, s, half, au_off, av_off, du, dv,
                     chis, n, out);
  return hipGetLastError();
}

static int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>(4096, (n + 255) / 256)); }

hipError_t launch_convert_msgs(const double* src, float* dst, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}
hipError_t launch_convert_msgs_back(const float* src, double* dst, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}
hipError_t launch_psi_to_f64(int dtype, const void* src, double* dst, long long n, hipStream_t s) {
  switch (dtype) {
    case ITNBP_F64: hipLaunchKernelGGL(k_to_f64<double>, dim3(grid_for(n)), dim3(256), 0, s, (const double*)src, dst, n); break;
    case ITNBP_F32: hipLaunchKernelGGL(k_to_f64<float>, dim3(grid_for(n)), dim3(256), 0, s, (const float*)src, dst, n); break;
    default: hipLaunchKernelGGL(k_to_f64<bf16_t>, dim3(grid_for(n)), dim3(256), 0, s, (const bf16_t*)src, dst, n); break;
  }
  return hipGetLastError();
}
hipError_t launch_scale_psi(int dtype, void* psi, const long long* offs, const int* sizes,
                            const double* scale, int nv, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  dim3 grid(64, nv);
  switch (dtype) {
    case ITNBP_F64: hipLaunchKernelGGL(k_scale_psi<double>, grid, dim3(256), 0, s, (double*)psi, offs, sizes, scale); break;
    case ITNBP_F32: hipLaunchKernelGGL(k_scale_psi<float>, grid, dim3(256), 0, s, (float*)psi, offs, sizes, scale); break;
    default: hipLaunchKernelGGL(k_scale_psi<bf16_t>, grid, dim3(256), 0, s, (bf16_t*)psi, offs, sizes, scale); break;
  }
  return hipGetLastError();
}
