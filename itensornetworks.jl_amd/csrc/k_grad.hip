// Hand-written adjoint of BP: gradient of E = sum_v <O_v> + sum_e <O_e> after
// T unrolled Jacobi sweeps from identity messages, with respect to every site
// tensor.  The reference has no autodiff; the forward semantics are those of
// updated_message + normalisation (src/caches/abstractbeliefpropagationcache.jl:225-239),
// expect (src/expect.jl:5-19) and the pair-environment RDM
// (test/test_belief_propagation.jl:64-91).
//
// Reverse pass, one workgroup per vertex u (deterministic, no atomics):
//   * a message u->v:  M = T/|T|  =>  Tbar = (Mbar - M <M, Mbar>)/|T|;
//   * an open contraction T[x, y] = sum_r Phi[x, r] psi[y, r] with
//     Phi = psi x_{a1} M_{a1} ... x_{an} M_{an}:
//       psibar_bra[y, r] += sum_x Tbar[x, y] Phi[x, r]
//       Phibar[x, r]      = sum_y Tbar[x, y] psi[y, r]
//       then for i = n..1: Mbar_{ai}[p, q] += <X_{i-1}[..p..], Xbar_i[..q..]>,
//                          Xbar_{i-1} = Xbar_i x_{ai} M_{ai}^T,
//       psibar_ket = Xbar_0.
// The cotangents of the messages flowing into u are produced only by u's own
// terms, so each workgroup owns its outputs.  Sweeps are checkpointed every c
// sweeps and recomputed segment by segment in the backward pass.
#include <cmath>
#include <vector>

#include "itnbp_ops.hpp"

namespace {

constexpr int GR_LDS_DOUBLES = 19200;  // 150 KiB of LDS per workgroup

struct GradVertex {
  int u;
  int nterm_energy;   // mode B: number of half-pair terms
  long long rho_bar;  // offset of rho_bar_u (d x d) in the cotangent buffer (-1: none)
  long long abar[ITNBP_MAXK];  // offsets of Abar for the half-pair term of bond axis a+1 (-1: none)
};

struct Geo {
  int ds, Aj, nj, lo_n, P, R;
};

__device__ inline Geo geo(const VDesc& V, int j, int s_open) {
  Geo g;
  g.ds = s_open ? V.dims[0] : 1;
  if (j > 0) {
    int A = 1;
    for (int a = 0; a < j; ++a) A *= V.dims[a];
    g.Aj = A;
    g.nj = V.dims[j];
  } else {
    g.Aj = V.D;
    g.nj = 1;
  }
  g.lo_n = g.Aj / g.ds;
  g.P = g.ds * g.nj;
  g.R = V.D / g.P;
  return g;
}

__device__ inline int offx(const Geo& g, int x) { return (x % g.ds) + g.Aj * (x / g.ds); }
__device__ inline int roff(const Geo& g, int r) { return g.ds * (r % g.lo_n) + g.Aj * g.nj * (r / g.lo_n); }

__device__ double gsum(double v, double* red) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
  __syncthreads();
  return s;
}

// Y = X x_a M (mode product, Y[..c..] = sum_x X[..x..] M[x + n c]); transpose: M^T
__device__ void mode_prod(const VDesc& V, int a, const double* X, double* Y, const double* M, bool transpose) {
  int A = 1;
  for (int b = 0; b < a; ++b) A *= V.dims[b];
  const int n = V.dims[a];
  for (int i = threadIdx.x; i < V.D; i += blockDim.x) {
    const int lo = i % A, rest = i / A, c = rest % n, hi = rest / n;
    const int base = lo + A * n * hi;
    double acc = 0;
    for (int x = 0; x < n; ++x) acc += X[base + A * x] * (transpose ? M[c + n * x] : M[x + n * c]);
    Y[i] = acc;
  }
  __syncthreads();
}

// C[x, y] = sum_r A[offx(x) + roff(r)] B[offx(y) + roff(r)] (open contraction between two tensors)
__device__ void open_contract(const Geo& g, const double* A, const double* B, double* C) {
  const int P2 = g.P * g.P;
  int G = 1;
  while (G * 2 * P2 <= (int)blockDim.x && G < 64) G *= 2;
  const int per = blockDim.x / G;
  for (int ob = 0; ob < P2; ob += per) {
    const int o = ob + threadIdx.x / G, gi = threadIdx.x % G;
    double acc = 0;
    if (o < P2) {
      const int ox = offx(g, o % g.P), oy = offx(g, o / g.P);
      for (int r = gi; r < g.R; r += G) {
        const int ro = roff(g, r);
        acc += A[ox + ro] * B[oy + ro];
      }
    }
    for (int m = G / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m);
    if (o < P2 && gi == 0) C[o] = acc;
  }
  __syncthreads();
}

// VJP of one open contraction of vertex V with cotangent Tbar (P x P in LDS):
// accumulates psibar (LDS, D) and Mbar of every absorbed axis (LDS, per axis chi^2).
// Xs: n+1 buffers of D (Xs[0] = psi); Xb0, Xb1: D each.
__device__ void vjp_term(const VDesc& V, const Geo& g, int nabs, const int* axes, const double* const* Ms,
                         double* const* Xs, const double* Tbar, double* psibar, double* const* Mbar, double* Xb0,
                         double* Xb1) {
  const double* Phi = Xs[nabs];
  const double* psi = Xs[0];
  // psibar_bra[y, r] += sum_x Tbar[x, y] Phi[x, r];  Phibar[x, r] = sum_y Tbar[x, y] psi[y, r]
  for (int idx = threadIdx.x; idx < V.D; idx += blockDim.x) {
    const int y = idx % g.P, r = idx / g.P;
    const int ro = roff(g, r);
    double a = 0, b = 0;
    for (int x = 0; x < g.P; ++x) {
      a += Tbar[x + g.P * y] * Phi[offx(g, x) + ro];
      b += Tbar[y + g.P * x] * psi[offx(g, x) + ro];
    }
    psibar[offx(g, y) + ro] += a;
    Xb0[offx(g, y) + ro] = b;
  }
  __syncthreads();
  double* cur = Xb0;
  double* oth = Xb1;
  for (int i = nabs; i >= 1; --i) {
    const int a = axes[i - 1];
    // Mbar[p, q] += sum X_{i-1}[..p..] Xbar_i[..q..]
    Geo ga = geo(V, a, 0);
    const int n = V.dims[a];
    double* mb = Mbar[i - 1];
    {
      const int P2 = n * n;
      int G = 1;
      while (G * 2 * P2 <= (int)blockDim.x && G < 64) G *= 2;
      const int per = blockDim.x / G;
      for (int ob = 0; ob < P2; ob += per) {
        const int o = ob + threadIdx.x / G, gi = threadIdx.x % G;
        double acc = 0;
        if (o < P2) {
          const int ox = offx(ga, o % n), oy = offx(ga, o / n);
          for (int r = gi; r < ga.R; r += G) {
            const int ro = roff(ga, r);
            acc += Xs[i - 1][ox + ro] * cur[oy + ro];
          }
        }
        for (int m = G / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m);
        if (o < P2 && gi == 0) mb[o] += acc;  // column-major: o = p + n q
      }
      __syncthreads();
    }
    mode_prod(V, a, cur, oth, Ms[i - 1], true);
    double* t = cur;
    cur = oth;
    oth = t;
  }
  for (int i = threadIdx.x; i < V.D; i += blockDim.x) psibar[i] += cur[i];
  __syncthreads();
}

// mode 0: backward through one sweep (messages out of u, normalisation);
// mode 1: energy terms (rho_bar_u and the half-pair Abar of every incident bond)
__global__ __launch_bounds__(256) void k_vertex_vjp(const VDesc* __restrict__ vds,
                                                    const GradVertex* __restrict__ gvs, int mode,
                                                    const double* __restrict__ psi,
                                                    const double* __restrict__ msg_prev,
                                                    const double* __restrict__ mbar_in,
                                                    double* __restrict__ mbar_out,
                                                    const long long* __restrict__ out_msg_off,
                                                    const long long* __restrict__ in_msg_off,
                                                    const double* __restrict__ cot,
                                                    double* __restrict__ grad, int normalize,
                                                    double* __restrict__ part, long long pstride) {
  extern __shared__ __align__(16) double sm[];
  __shared__ double red[8];
  const GradVertex gv = gvs[blockIdx.x];
  const int u = gv.u;
  const VDesc& V = vds[u];  // by reference: a copy indexed at run time lands in scratch
  const int D = V.D, k = V.k;
  const int tid = threadIdx.x;
  // LDS: Xs[0..k] (k+1)D, Xb0, Xb1, psibar: 3D, Tb (P^2 <= 4096), Mbar k x chi^2, T
  double* Xs[ITNBP_MAXR + 1];
  for (int i = 0; i <= k; ++i) Xs[i] = sm + (long long)i * D;
  double* Xb0 = sm + (long long)(k + 1) * D;
  double* Xb1 = Xb0 + D;
  double* psibar = Xb1 + D;
  double* Mb = psibar + D;  // per bond axis a (1..k): chi_a^2, at Mb + mb_off[a]
  int mb_off[ITNBP_MAXR];
  int off = 0;
  for (int a = 1; a <= k; ++a) {
    mb_off[a] = off;
    off += V.dims[a] * V.dims[a];
  }
  double* Tb = Mb + off;  // up to 4096
  double* Tm = Tb + 4096;
  for (int i = tid; i < D; i += blockDim.x) {
    Xs[0][i] = psi[V.psi_off + i];
    psibar[i] = 0;
  }
  for (int i = tid; i < off; i += blockDim.x) Mb[i] = 0;
  __syncthreads();
  const int nterms = mode == 0 ? k : 1 + k;
  // with part: workgroup (x, y) takes term y of vertex x only and writes its
  // contributions (psibar, then Mbar of every bond) to part, zeros for a term that
  // does not exist; k_vjp_reduce sums the terms
  const int t0 = part ? (int)blockIdx.y : 0;
  const int t1 = part ? min((int)blockIdx.y + 1, nterms) : nterms;
  for (int term = t0; term < t1; ++term) {
    int j, s_open;
    if (mode == 0) {
      j = term + 1;
      s_open = 0;
    } else if (term == 0) {
      if (gv.rho_bar < 0) continue;
      j = 0;
      s_open = 1;
    } else {
      j = term;
      if (gv.abar[j - 1] < 0) continue;
      s_open = 1;
    }
    int axes[ITNBP_MAXK];
    const double* Ms[ITNBP_MAXK];
    double* Mbars[ITNBP_MAXK];
    int nabs = 0;
    for (int a = 1; a <= k; ++a) {
      if (a == j) continue;
      axes[nabs] = a;
      Ms[nabs] = msg_prev + V.in_msg[a];
      Mbars[nabs] = Mb + mb_off[a];
      ++nabs;
    }
    for (int i = 1; i <= nabs; ++i) mode_prod(V, axes[i - 1], Xs[i - 1], Xs[i], Ms[i - 1], false);
    const Geo g = geo(V, j, s_open);
    const int P2 = g.P * g.P;
    if (mode == 0) {
      // T, its norm, M = T/|T|, Tbar = (Mbar - M <M, Mbar>)/|T|
      open_contract(g, Xs[nabs], Xs[0], Tm);
      const double* mb_in = mbar_in + out_msg_off[(long long)blockIdx.x * ITNBP_MAXK + term];
      double n2 = 0;
      for (int o = tid; o < P2; o += blockDim.x) n2 += Tm[o] * Tm[o];
      n2 = gsum(n2, red);
      const double nrm = sqrt(n2);
      if (normalize && nrm != 0.0) {
        double dot = 0;
        for (int o = tid; o < P2; o += blockDim.x) dot += (Tm[o] / nrm) * mb_in[o];
        dot = gsum(dot, red);
        for (int o = tid; o < P2; o += blockDim.x) Tb[o] = (mb_in[o] - (Tm[o] / nrm) * dot) / nrm;
      } else {
        for (int o = tid; o < P2; o += blockDim.x) Tb[o] = mb_in[o];
      }
    } else {
      const double* src = term == 0 ? cot + gv.rho_bar : cot + gv.abar[j - 1];
      for (int o = tid; o < P2; o += blockDim.x) Tb[o] = src[o];
    }
    __syncthreads();
    vjp_term(V, g, nabs, axes, Ms, Xs, Tb, psibar, Mbars, Xb0, Xb1);
  }
  // outputs
  if (part) {
    double* pp = part + ((long long)blockIdx.x * gridDim.y + blockIdx.y) * pstride;
    for (int i = tid; i < D + off; i += blockDim.x) pp[i] = i < D ? psibar[i] : Mb[i - D];
    return;
  }
  for (int i = tid; i < D; i += blockDim.x) grad[V.psi_off + i] += psibar[i];
  for (int a = 1; a <= k; ++a) {
    const int n2 = V.dims[a] * V.dims[a];
    double* dst = mbar_out + in_msg_off[(long long)blockIdx.x * ITNBP_MAXK + a - 1];
    for (int i = tid; i < n2; i += blockDim.x) dst[i] = Mb[mb_off[a] + i];
  }
}

// Sum of the per-term partials of k_vertex_vjp (terms in order): grad += psibar,
// Mbar of every incoming message = the sum.  grid (vertices), 256 threads.
__global__ __launch_bounds__(256) void k_vjp_reduce(const VDesc* __restrict__ vds,
                                                    const GradVertex* __restrict__ gvs, int mode,
                                                    const double* __restrict__ part, long long pstride,
                                                    int nty, double* __restrict__ mbar_out,
                                                    const long long* __restrict__ in_msg_off,
                                                    double* __restrict__ grad) {
  const GradVertex gv = gvs[blockIdx.x];
  const VDesc& V = vds[gv.u];  // by reference: a copy indexed at run time lands in scratch
  const int D = V.D, k = V.k;
  const int nterms = mode == 0 ? k : 1 + k;
  const double* pp = part + (long long)blockIdx.x * nty * pstride;
  for (int i = threadIdx.x; i < D; i += blockDim.x) {
    double acc = 0;
    for (int t = 0; t < nterms; ++t) acc += pp[(long long)t * pstride + i];
    grad[V.psi_off + i] += acc;
  }
  int off = 0;
  for (int a = 1; a <= k; ++a) {
    const int n2 = V.dims[a] * V.dims[a];
    double* dst = mbar_out + in_msg_off[(long long)blockIdx.x * ITNBP_MAXK + a - 1];
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      double acc = 0;
      for (int t = 0; t < nterms; ++t) acc += pp[(long long)t * pstride + D + off + i];
      dst[i] = acc;
    }
    off += n2;
  }
}

// per-vertex rho_bar and energy terms
__global__ void k_rho_bar1(const double* __restrict__ rho, const double* __restrict__ ops,
                           const long long* __restrict__ off, const int* __restrict__ dims, int n,
                           double* __restrict__ rbar, double* __restrict__ eterm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int d = dims[i];
  const double* r = rho + off[i];
  const double* O = ops + off[i];
  double num = 0, den = 0;
  for (int s = 0; s < d; ++s) {
    den += r[s + d * s];
    for (int sp = 0; sp < d; ++sp) num += O[sp + d * s] * r[s + d * sp];
  }
  const double ev = num / den;
  eterm[i] = ev;
  for (int s = 0; s < d; ++s)
    for (int sp = 0; sp < d; ++sp) rbar[off[i] + s + d * sp] = (O[sp + d * s] - (s == sp ? ev : 0.0)) / den;
}

// per-edge: rho_bar_e and Abar_u, Abar_v from the halves
__global__ void k_pair_bar(const double* __restrict__ half, const long long* __restrict__ au_off,
                           const long long* __restrict__ av_off, const double* __restrict__ rho,
                           const double* __restrict__ ops, const long long* __restrict__ roff2,
                           const int* __restrict__ dus, const int* __restrict__ dvs,
                           const int* __restrict__ chis, double* __restrict__ abar_u,
                           double* __restrict__ abar_v, double* __restrict__ eterm) {
  __shared__ double rb[256];
  __shared__ double evs;
  const int p = blockIdx.x;
  const int du = dus[p], dv = dvs[p], chi = chis[p], dd = du * dv;
  const double* r = rho + roff2[p];
  const double* O = ops + roff2[p];
  if (threadIdx.x == 0) {
    double num = 0, den = 0;
    for (int s = 0; s < dd; ++s) {
      den += r[s + dd * s];
      for (int sp = 0; sp < dd; ++sp) num += O[sp + dd * s] * r[s + dd * sp];
    }
    evs = num / den;
    eterm[p] = evs;
    for (int s = 0; s < dd; ++s)
      for (int sp = 0; sp < dd; ++sp) rb[s + dd * sp] = (O[sp + dd * s] - (s == sp ? evs : 0.0)) / den;
  }
  __syncthreads();
  const double* Au = half + au_off[p];
  const double* Av = half + av_off[p];
  const int lu = du * chi, lv = dv * chi;
  double* bu = abar_u + au_off[p];
  double* bv = abar_v + av_off[p];
  for (int idx = threadIdx.x; idx < lu * lu; idx += blockDim.x) {
    const int row = idx % lu, col = idx / lu;
    const int su = row % du, b = row / du, su2 = col % du, b2 = col / du;
    double acc = 0;
    for (int sv = 0; sv < dv; ++sv)
      for (int sv2 = 0; sv2 < dv; ++sv2)
        acc += rb[(su + du * sv) + dd * (su2 + du * sv2)] * Av[(sv + dv * b) + (long long)lv * (sv2 + dv * b2)];
    bu[idx] = acc;
  }
  for (int idx = threadIdx.x; idx < lv * lv; idx += blockDim.x) {
    const int row = idx % lv, col = idx / lv;
    const int sv = row % dv, b = row / dv, sv2 = col % dv, b2 = col / dv;
    double acc = 0;
    for (int su = 0; su < du; ++su)
      for (int su2 = 0; su2 < du; ++su2)
        acc += rb[(su + du * sv) + dd * (su2 + du * sv2)] * Au[(su + du * b) + (long long)lu * (su2 + du * b2)];
    bv[idx] = acc;
  }
}

}  // namespace

// forward sweeps helper (itnbp_api.hip)
int sweep_phase_public(itnbp_net* net, int nsweeps, double* trace = nullptr, long long trace_stride = 0,
                       int trace_every = 1);

namespace {

// The gradient of the energy (ops1, ops2) after nsweeps sweeps, or with
// mbar_seed the pullback of a given cotangent of the final messages M^(T)
// through the sweeps (no energy terms).
int energy_grad_impl(itnbp_net* net, int nsweeps, int checkpoint, const double* ops1, const double* ops2,
                     const double* mbar_seed, double* energy, double* grad) {
  if (!net || nsweeps < 0 || !energy || !grad) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  if (net->dtype != ITNBP_F64) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "energy_grad needs FP64 site tensors");
  const int nv = net->nv, ne = net->ne;
  const int c = checkpoint <= 0 ? 1 : checkpoint;
  // LDS check per vertex
  for (int v = 0; v < nv; ++v) {
    const VDesc& V = net->vdesc[v];
    long long need = (long long)(V.k + 4) * V.D + 4096 + 4096;
    for (int a = 1; a <= V.k; ++a) need += (long long)V.dims[a] * V.dims[a];
    if (need > GR_LDS_DOUBLES)
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "energy_grad: site tensor too large for the LDS adjoint kernel");
    if (V.dims[0] * (V.k > 0 ? 1 : 1) > 64) return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "energy_grad: d too large");
  }
  for (int e = 0; e < ne; ++e) {
    const int du = net->phys[net->edges[2 * e]], dv = net->phys[net->edges[2 * e + 1]];
    if (du * dv > 16 || du * net->chi[e] > 64 || dv * net->chi[e] > 64)
      return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "energy_grad: pair too large");
  }
  const long long MT = net->msg_total;
  const int ncheck = nsweeps / c + 1;
  DevBuf ck, seg, mbar, cot, tables, gradbuf, vpart;
  auto cleanup = [&]() {
    free_buf(vpart);
    free_buf(ck);
    free_buf(seg);
    free_buf(mbar);
    free_buf(cot);
    free_buf(tables);
    free_buf(gradbuf);
  };
#define GR_CHK(call)             \
  do {                           \
    int _r = (call);             \
    if (_r != ITNBP_OK) {        \
      cleanup();                 \
      return _r;                 \
    }                            \
  } while (0)
#define GR_HIP(call)                                                                           \
  do {                                                                                         \
    hipError_t _e = (call);                                                                    \
    if (_e != hipSuccess) {                                                                    \
      cleanup();                                                                               \
      return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    }                                                                                          \
  } while (0)
  GR_CHK(ensure_buf(ctx, ck, (size_t)ncheck * MT * 8));
  GR_CHK(ensure_buf(ctx, seg, (size_t)std::max(1, c) * MT * 8));
  GR_CHK(ensure_buf(ctx, mbar, (size_t)2 * MT * 8));
  double* d_ck = (double*)ck.p;
  double* d_seg = (double*)seg.p;
  // forward: identity messages, nsweeps Jacobi sweeps, checkpoints every c sweeps
  GR_CHK(itnbp_init_identity_messages(net));
  GR_HIP(hipMemcpyAsync(d_ck, net->d_msg[net->cur], MT * 8, hipMemcpyDeviceToDevice, ctx->stream));
  // (one multi-sweep call; the checkpoint after sweep t, t % c == 0, to d_ck + (t / c) MT)
  if (nsweeps > 0) GR_CHK(sweep_phase_public(net, nsweeps, d_ck, MT, c));
  // energy forward: rho_v (d x d) and halves A for every edge, rho_e
  std::vector<long long> r1off(nv), aoff_u(ne), aoff_v(ne), r2off(ne);
  std::vector<int> dims1(nv), dus(ne), dvs(ne), chis(ne);
  long long n1 = 0, nh = 0, n2 = 0;
  for (int v = 0; v < nv; ++v) {
    r1off[v] = n1;
    dims1[v] = net->phys[v];
    n1 += (long long)net->phys[v] * net->phys[v];
  }
  for (int e = 0; e < ne; ++e) {
    const int u = net->edges[2 * e], v = net->edges[2 * e + 1];
    dus[e] = net->phys[u];
    dvs[e] = net->phys[v];
    chis[e] = net->chi[e];
    aoff_u[e] = nh;
    nh += (long long)(dus[e] * chis[e]) * (dus[e] * chis[e]);
    aoff_v[e] = nh;
    nh += (long long)(dvs[e] * chis[e]) * (dvs[e] * chis[e]);
    r2off[e] = n2;
    n2 += (long long)(dus[e] * dvs[e]) * (dus[e] * dvs[e]);
  }
  // buffers: rho1 | rho2 | half | rbar1 | abar | eterms | ops1 | ops2
  const long long o_rho1 = 0, o_rho2 = o_rho1 + n1, o_half = o_rho2 + n2, o_rbar1 = o_half + nh,
                  o_abar = o_rbar1 + n1, o_et = o_abar + nh, o_op1 = o_et + nv + ne, o_op2 = o_op1 + n1,
                  o_end = o_op2 + n2;
  GR_CHK(ensure_buf(ctx, cot, (size_t)std::max<long long>(1, o_end) * 8));
  double* C = (double*)cot.p;
  GR_HIP(hipMemsetAsync(C, 0, o_end * 8, ctx->stream));
  if (ops1) GR_HIP(hipMemcpyAsync(C + o_op1, ops1, n1 * 8, hipMemcpyHostToDevice, ctx->stream));
  if (ops2) GR_HIP(hipMemcpyAsync(C + o_op2, ops2, n2 * 8, hipMemcpyHostToDevice, ctx->stream));
  {
    std::vector<OcTask> tasks;
    for (int v = 0; ops1 && v < nv; ++v) {
      OcTask t;
      t.v = v;
      t.j = 0;
      t.s_open = 1;
      t.absorb_mask = all_bonds_mask(net, v);
      t.out_off = o_rho1 + r1off[v];
      t.old_off = -1;
      t.normalize = 0;
      t.slot = v;
      tasks.push_back(t);
    }
    for (int e = 0; ops2 && e < ne; ++e)
      for (int side = 0; side < 2; ++side) {
        const int x = net->edges[2 * e + side];
        OcTask t;
        t.v = x;
        t.j = axis_of(net, x, e);
        t.s_open = 1;
        t.absorb_mask = all_bonds_mask(net, x) & ~(1 << t.j);
        t.out_off = o_half + (side == 0 ? aoff_u[e] : aoff_v[e]);
        t.old_off = -1;
        t.normalize = 0;
        t.slot = nv + 2 * e + side;
        tasks.push_back(t);
      }
    GR_CHK(ensure_buf(ctx, net->stats, sizeof(MsgStat) * (size_t)std::max(1, nv + 2 * ne)));
    GR_CHK(run_tasks(net, tasks, 1, net->d_psi, net->dtype, net->d_msg[net->cur], C, (MsgStat*)net->stats.p,
                     net->tasks));
  }
  // tables for the pair kernels and the vertex kernel
  std::vector<GradVertex> gvs(nv);
  std::vector<long long> outoff((size_t)nv * ITNBP_MAXK, 0), inoff((size_t)nv * ITNBP_MAXK, 0);
  for (int v = 0; v < nv; ++v) {
    GradVertex& g = gvs[v];
    g.u = v;
    g.rho_bar = ops1 ? o_rbar1 + r1off[v] : -1;
    g.nterm_energy = 0;
    const int k = net->vdesc[v].k;
    for (int a = 0; a < ITNBP_MAXK; ++a) g.abar[a] = -1;
    for (int a = 1; a <= k; ++a) {
      const int e = net->inc_edge[net->inc_ptr[v] + a - 1];
      const int w = net->edges[2 * e] == v ? net->edges[2 * e + 1] : net->edges[2 * e];
      if (ops2) g.abar[a - 1] = o_abar + (net->edges[2 * e] == v ? aoff_u[e] : aoff_v[e]);
      outoff[(size_t)v * ITNBP_MAXK + a - 1] = net->msg_off[msg_id_into(net, e, w)];
      inoff[(size_t)v * ITNBP_MAXK + a - 1] = net->msg_off[msg_id_into(net, e, v)];
    }
  }
  const size_t tb_bytes = nv * sizeof(GradVertex) + 2 * outoff.size() * 8 + ne * (8 * 3 + 4 * 3) + nv * (8 + 4) + 64;
  GR_CHK(ensure_buf(ctx, tables, tb_bytes));
  char* tb = (char*)tables.p;
  size_t to = 0;
  auto put = [&](const void* src, size_t bytes) -> void* {
    void* dst = tb + to;
    (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    to += (bytes + 15) / 16 * 16;
    return dst;
  };
  const GradVertex* d_gv = (const GradVertex*)put(gvs.data(), nv * sizeof(GradVertex));
  const long long* d_out = (const long long*)put(outoff.data(), outoff.size() * 8);
  const long long* d_in = (const long long*)put(inoff.data(), inoff.size() * 8);
  std::vector<long long> au_abs(ne), av_abs(ne);
  for (int e = 0; e < ne; ++e) {
    au_abs[e] = aoff_u[e];
    av_abs[e] = aoff_v[e];
  }
  const long long* d_au = ne ? (const long long*)put(au_abs.data(), ne * 8) : nullptr;
  const long long* d_av = ne ? (const long long*)put(av_abs.data(), ne * 8) : nullptr;
  const long long* d_r2 = ne ? (const long long*)put(r2off.data(), ne * 8) : nullptr;
  const int* d_du = ne ? (const int*)put(dus.data(), ne * 4) : nullptr;
  const int* d_dv = ne ? (const int*)put(dvs.data(), ne * 4) : nullptr;
  const int* d_chi = ne ? (const int*)put(chis.data(), ne * 4) : nullptr;
  const long long* d_r1 = (const long long*)put(r1off.data(), nv * 8);
  const int* d_d1 = (const int*)put(dims1.data(), nv * 4);
  GR_HIP(hipStreamSynchronize(ctx->stream));
  if (to > tb_bytes) {
    cleanup();
    return itnbp_fail(ctx, ITNBP_ERR_ARG, "energy_grad: table overflow");
  }
  // rho_e = sum halves
  if (ops2 && ne > 0) {
    // the combine kernel writes pair p at p * (du dv)^2: uniform d only
    for (int e = 1; e < ne; ++e)
      if (dus[e] != dus[0] || dvs[e] != dvs[0]) {
        cleanup();
        return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "energy_grad: two-site terms need uniform d");
      }
    GR_HIP(launch_rdm2_combine(C + o_half, d_au, d_av, d_du, d_dv, d_chi, ne, C + o_rho2, ctx->stream));
    hipLaunchKernelGGL(k_pair_bar, dim3(ne), dim3(256), 0, ctx->stream, C + o_half, d_au, d_av, C + o_rho2,
                       C + o_op2, d_r2, d_du, d_dv, d_chi, C + o_abar, C + o_abar, C + o_et + nv);
    GR_HIP(hipGetLastError());
  }
  if (ops1) {
    hipLaunchKernelGGL(k_rho_bar1, dim3((nv + 255) / 256), dim3(256), 0, ctx->stream, C + o_rho1, C + o_op1,
                       d_r1, d_d1, nv, C + o_rbar1, C + o_et);
    GR_HIP(hipGetLastError());
  }
  std::vector<double> et(nv + ne, 0.0);
  GR_HIP(hipMemcpyAsync(et.data(), C + o_et, (nv + ne) * 8, hipMemcpyDeviceToHost, ctx->stream));
  GR_HIP(hipStreamSynchronize(ctx->stream));
  long double E = 0;
  if (ops1)
    for (int v = 0; v < nv; ++v) E += et[v];
  if (ops2)
    for (int e = 0; e < ne; ++e) E += et[nv + e];
  *energy = (double)E;
  // backward
  const size_t lds = (size_t)GR_LDS_DOUBLES * 8;
  GR_HIP(hipFuncSetAttribute((const void*)k_vertex_vjp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  GR_CHK(ensure_buf(ctx, gradbuf, net->psi_total * 8));
  double* d_grad = (double*)gradbuf.p;
  GR_HIP(hipMemsetAsync(d_grad, 0, net->psi_total * 8, ctx->stream));
  double* mb[2] = {(double*)mbar.p, (double*)mbar.p + MT};
  // energy terms at the final messages M^(T)
  // one workgroup per (vertex, term), then the fixed-order sum of the terms
  int maxk = 0;
  long long pstride = 1;
  for (int v = 0; v < nv; ++v) {
    const VDesc& V = net->vdesc[v];
    long long w = V.D;
    for (int a = 1; a <= V.k; ++a) w += (long long)V.dims[a] * V.dims[a];
    maxk = std::max(maxk, V.k);
    pstride = std::max(pstride, w);
  }
  const int nty = maxk + 1;
  GR_CHK(ensure_buf(ctx, vpart, (size_t)std::max(1, nv) * nty * pstride * 8));
  double* d_part = (double*)vpart.p;
  hipLaunchKernelGGL(k_vertex_vjp, dim3(nv, nty), dim3(256), lds, ctx->stream, net->d_vdesc, d_gv, 1,
                     (const double*)net->d_psi, (const double*)net->d_msg[net->cur], (const double*)nullptr, mb[0],
                     d_out, d_in, (const double*)C, d_grad, 1, d_part, pstride);
  GR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_vjp_reduce, dim3(nv), dim3(256), 0, ctx->stream, net->d_vdesc, d_gv, 1,
                     (const double*)d_part, pstride, nty, mb[0], d_in, d_grad);
  GR_HIP(hipGetLastError());
  // a given cotangent of M^(T) replaces the (empty) energy terms' one
  if (mbar_seed) GR_HIP(hipMemcpyAsync(mb[0], mbar_seed, MT * 8, hipMemcpyHostToDevice, ctx->stream));
  int cb = 0;  // mb[cb] holds Mbar^(t)
  // segments in reverse: segment s covers t-1 in [s c, s c + c - 1]
  for (int s = (nsweeps - 1) / c; s >= 0 && nsweeps > 0; --s) {
    const int base = s * c;
    const int top = std::min(nsweeps - 1, base + c - 1);
    // recompute M^(base+1..top) from the checkpoint into seg
    GR_HIP(hipMemcpyAsync(net->d_msg[net->cur], d_ck + (long long)s * MT, MT * 8, hipMemcpyDeviceToDevice,
                          ctx->stream));
    // M^(t) for t = base + 1 .. top to d_seg + (t - base) MT, in one call
    if (top > base) GR_CHK(sweep_phase_public(net, top - base, d_seg, MT, 1));
    GR_HIP(hipMemcpyAsync(d_seg, d_ck + (long long)s * MT, MT * 8, hipMemcpyDeviceToDevice, ctx->stream));
    for (int tm1 = top; tm1 >= base; --tm1) {
      const double* prev = d_seg + (long long)(tm1 - base) * MT;
      if (maxk == 0) continue;  // no messages: nothing flows back through the sweeps
      hipLaunchKernelGGL(k_vertex_vjp, dim3(nv, maxk), dim3(256), lds, ctx->stream, net->d_vdesc, d_gv, 0,
                         (const double*)net->d_psi, prev, (const double*)mb[cb], mb[1 - cb], d_out, d_in,
                         (const double*)C, d_grad, 1, d_part, pstride);
      GR_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_vjp_reduce, dim3(nv), dim3(256), 0, ctx->stream, net->d_vdesc, d_gv, 0,
                         (const double*)d_part, pstride, maxk, mb[1 - cb], d_in, d_grad);
      GR_HIP(hipGetLastError());
      cb = 1 - cb;
    }
  }
  GR_HIP(hipMemcpyAsync(grad, d_grad, net->psi_total * 8, hipMemcpyDeviceToHost, ctx->stream));
  GR_HIP(hipStreamSynchronize(ctx->stream));
  // leave the cache holding the final forward messages M^(T)
  GR_HIP(hipMemcpyAsync(net->d_msg[net->cur], d_ck + (long long)(nsweeps / c) * MT, MT * 8,
                        hipMemcpyDeviceToDevice, ctx->stream));
  if (nsweeps % c != 0) GR_CHK(sweep_phase_public(net, nsweeps % c));
  GR_HIP(hipStreamSynchronize(ctx->stream));
  cleanup();
  return ITNBP_OK;
#undef GR_CHK
#undef GR_HIP
}

}  // namespace

extern "C" int itnbp_energy_grad(itnbp_net* net, int nsweeps, int checkpoint, const double* ops1,
                                 const double* ops2, double* energy, double* grad) {
  if (!net) return ITNBP_ERR_ARG;
  return itnbp_guarded(net->ctx, [&]() -> int {
    return energy_grad_impl(net, nsweeps, checkpoint, ops1, ops2, nullptr, energy, grad);
  });
}

extern "C" int itnbp_sweeps_vjp(itnbp_net* net, int nsweeps, int checkpoint, const double* mbar, double* grad) {
  if (!net || !mbar) return ITNBP_ERR_ARG;
  double unused = 0;
  return itnbp_guarded(net->ctx, [&]() -> int {
    return energy_grad_impl(net, nsweeps, checkpoint, nullptr, nullptr, mbar, &unused, grad);
  });
}
