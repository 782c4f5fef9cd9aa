// Two-layer belief propagation: the BP hot path on a network whose bra and ket
// layers are different states (and may have different bond dimensions), real
// or complex, in FP64:
//   - inner(phi, psi; alg="bp") runs BP on the LinearFormNetwork <phi|psi>
//     (src/inner.jl:128-160, src/formnetworks/linearformnetwork.jl:15-25) with
//     one {bra, ket} partition per site;
//   - inner(phi, A, psi; alg="bp") runs it on the BilinearFormNetwork
//     <phi|A|psi> (src/formnetworks/bilinearformnetwork.jl:23-42); merging the
//     operator's and the ket's link indices on every edge makes it the
//     two-layer network <phi|A psi>, whose ket site tensors k_layer_fold builds;
//   - with phi = psi it is the norm network of a complex state.
// Messages are M[b_ket, b_bra] (chi_ket x chi_bra, column-major).  A message
// update (src/caches/abstractbeliefpropagationcache.jl:225-239) absorbs the
// incoming messages into the ket site tensor one bond at a time (each turns a
// ket bond into a bra bond) and contracts the result with conj(bra) over every
// index but the open bond; message_diff is 1 - |<a/|a|, b/|b|>|^2 (:32-36).
// The same kernel with the physical index open gives the site's density
// matrix, and with nothing open the vertex scalar (beliefpropagationcache.jl:107-113).
#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "itnbp_graph.hpp"
#include "itnbp_internal.hpp"
#include "itnbp_ops.hpp"

namespace {

struct cplx {
  double x, y;
};
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return cplx{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return cplx{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) { return cplx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline cplx operator*(cplx a, double s) { return cplx{a.x * s, a.y * s}; }
__host__ __device__ inline cplx& operator+=(cplx& a, cplx b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}
__host__ __device__ inline double conjv(double a) { return a; }
__host__ __device__ inline cplx conjv(cplx a) { return cplx{a.x, -a.y}; }
__host__ __device__ inline double abs2v(double a) { return a * a; }
__host__ __device__ inline double abs2v(cplx a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline bool finitev(double a) { return isfinite(a); }
__host__ __device__ inline bool finitev(cplx a) { return isfinite(a.x) && isfinite(a.y); }
__host__ __device__ inline double rev(double a) { return a; }
__host__ __device__ inline double rev(cplx a) { return a.x; }
__host__ __device__ inline double imv(double) { return 0.0; }
__host__ __device__ inline double imv(cplx a) { return a.y; }
template <typename T>
__host__ __device__ inline T zerov() {
  return T{};
}

__device__ inline double shfl_xor_v(double v, int m) { return __shfl_xor(v, m); }
__device__ inline cplx shfl_xor_v(cplx v, int m) { return cplx{__shfl_xor(v.x, m), __shfl_xor(v.y, m)}; }

// sum over the block (256 threads); red holds 4 doubles
__device__ double block_sum_d(double v, double* red) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

constexpr int LY_THREADS = 256;

// one site of the two-layer network
struct LVDesc {
  int k, pad;
  int kd[ITNBP_MAXR];  // ket dims [d, chi_ket of every bond]
  int bd[ITNBP_MAXR];  // bra dims [d, chi_bra of every bond]
  long long ket_off, bra_off;
  long long in_msg[ITNBP_MAXR];  // offset of the message into the site along bond axis a
  long long Dk;                  // ket elements
};

// j >= 1: the message out along bond axis j (to out_off, compared with old_off,
// statistics to slot); j = 0: the density matrix R[s_ket, s_bra]; j = -1: the
// vertex scalar
struct LTask {
  int v, j, slot, normalize;
  long long out_off, old_off;
};

// grid (tasks), 256 threads; scratch: three areas of sstride elements per workgroup
template <typename T>
__global__ __launch_bounds__(LY_THREADS) void k_layer_task(const LVDesc* __restrict__ vds,
                                                          const LTask* __restrict__ tasks,
                                                          const T* __restrict__ ket, const T* __restrict__ bra,
                                                          const T* msg_in, T* out, MsgStat* __restrict__ stats,
                                                          T* __restrict__ scratch, long long sstride) {
  __shared__ double red[4];
  const LTask t = tasks[blockIdx.x];
  const LVDesc V = vds[t.v];
  const int k = V.k, tid = threadIdx.x;
  int dims[ITNBP_MAXR];
  for (int a = 0; a <= k; ++a) dims[a] = V.kd[a];
  long long D = V.Dk;
  T* area = scratch + (long long)blockIdx.x * 3 * sstride;
  const T* X = ket + V.ket_off;
  int cur = -1;
  // 1. absorb every incoming message but the open bond's: axis a goes from the
  //    ket dimension to the bra dimension
  for (int a = 1; a <= k; ++a) {
    if (a == t.j) continue;
    const int nk = V.kd[a], nb = V.bd[a];
    const T* M = msg_in + V.in_msg[a];
    long long A = 1;
    for (int b = 0; b < a; ++b) A *= dims[b];
    const long long B = D / (A * nk);
    const long long Dn = A * nb * B;
    T* Y = area + (long long)((cur + 1) & 1) * sstride;
    for (long long i = tid; i < Dn; i += LY_THREADS) {
      const long long lo = i % A, rest = i / A;
      const int c = (int)(rest % nb);
      const long long hi = rest / nb;
      const long long base = lo + A * nk * hi;
      T acc = zerov<T>();
      for (int x = 0; x < nk; ++x) acc += X[base + A * x] * M[x + (long long)nk * c];
      Y[i] = acc;
    }
    __threadfence_block();
    __syncthreads();
    X = Y;
    cur = (cur + 1) & 1;
    dims[a] = nb;
    D = Dn;
  }
  // 2. contract with conj(bra) over every index but the open one:
  //    T[o] = sum_r X[offx(row) + roff(r)] conj(bra[offy(col) + roff(r)])
  const T* P = bra + V.bra_off;
  long long Aj, nrow, ncol;
  if (t.j >= 1) {
    Aj = 1;
    for (int b = 0; b < t.j; ++b) Aj *= dims[b];
    nrow = V.kd[t.j];
    ncol = V.bd[t.j];
  } else if (t.j == 0) {
    Aj = 1;
    nrow = ncol = V.kd[0];
  } else {
    Aj = 1;
    nrow = ncol = 1;
  }
  const long long R = D / nrow;  // summed elements per output
  const int P2 = (int)(nrow * ncol);
  T* Tb = area + 2 * sstride;
  int G = 1;
  while (G * 2 * P2 <= LY_THREADS && G < 64) G *= 2;
  const int per = LY_THREADS / G;
  for (int ob = 0; ob < P2; ob += per) {
    const int o = ob + tid / G, gi = tid % G;
    T acc = zerov<T>();
    if (o < P2) {
      const long long row = o % nrow, col = o / nrow;
      for (long long r = gi; r < R; r += G) {
        const long long lo = r % Aj, hi = r / Aj;
        acc += X[lo + Aj * (row + nrow * hi)] * conjv(P[lo + Aj * (col + ncol * hi)]);
      }
    }
    for (int m = G / 2; m > 0; m >>= 1) acc += shfl_xor_v(acc, m);
    if (o < P2 && gi == 0) Tb[o] = acc;
  }
  __threadfence_block();
  __syncthreads();
  if (t.j < 1) {
    // density matrix or scalar: raw
    for (int o = tid; o < P2; o += LY_THREADS) out[t.out_off + o] = Tb[o];
    return;
  }
  // 3. message: normalise, compare with the previous message, write
  double n2 = 0;
  for (int o = tid; o < P2; o += LY_THREADS) n2 += abs2v(Tb[o]);
  n2 = block_sum_d(n2, red);
  const double nrm = sqrt(n2);
  const double sc = (t.normalize && nrm != 0.0) ? 1.0 / nrm : 1.0;
  double dre = 0, dim = 0, o2 = 0, d2 = 0, bad = 0;
  for (int o = tid; o < P2; o += LY_THREADS) {
    const T nv = Tb[o] * sc;
    if (!finitev(nv)) bad = 1;
    if (t.old_off >= 0) {
      const T ov = msg_in[t.old_off + o];
      const T pr = conjv(ov) * nv;
      dre += rev(pr);
      dim += imv(pr);
      o2 += abs2v(ov);
      d2 += abs2v(nv - ov);
    }
  }
  // every thread has read the old message before it is overwritten
  __syncthreads();
  for (int o = tid; o < P2; o += LY_THREADS) out[t.out_off + o] = Tb[o] * sc;
  dre = block_sum_d(dre, red);
  dim = block_sum_d(dim, red);
  o2 = block_sum_d(o2, red);
  d2 = block_sum_d(d2, red);
  bad = block_sum_d(bad, red);
  if (tid == 0 && t.slot >= 0) {
    const double nn2 = n2 * sc * sc;
    MsgStat st;
    st.diff = t.old_off >= 0 ? 1.0 - (dre * dre + dim * dim) / (o2 * nn2) : 0.0;
    st.res2 = t.old_off >= 0 ? (nn2 > 0 ? d2 / nn2 : 0.0) : 0.0;
    st.norm = nrm;
    st.bad = (bad > 0 || !isfinite(nrm)) ? 1.0 : 0.0;
    stats[t.slot] = st;
  }
}

// Edge scalars sum_{b, b'} M_{u->v}[b, b'] M_{v->u}[b, b']
// (beliefpropagationcache.jl:115-119): one wave per edge.
template <typename T>
__global__ __launch_bounds__(64) void k_layer_edge_scalar(const long long* __restrict__ moff, const int* __restrict__ msize,
                                                         const T* __restrict__ msg, T* __restrict__ out) {
  const int e = blockIdx.x, l = threadIdx.x;
  const T* a = msg + moff[2 * e];
  const T* b = msg + moff[2 * e + 1];
  T acc = zerov<T>();
  for (int i = l; i < msize[e]; i += 64) acc += a[i] * b[i];
  for (int m = 32; m > 0; m >>= 1) acc += shfl_xor_v(acc, m);
  if (l == 0) out[e] = acc;
}

// Ket of <phi|A psi>: out_v[s', c_1 .. c_k] = sum_s A_v[s', s, a_1 .. a_k] psi_v[s, b_1 .. b_k]
// with c = b + chi_psi a on every bond.  One thread per output element.
struct FoldDesc {
  int k, d;
  int pchi[ITNBP_MAXK], achi[ITNBP_MAXK];
  long long psi_off, op_off, out_off, Dout;
};

template <typename T>
__global__ __launch_bounds__(256) void k_layer_fold(const FoldDesc* __restrict__ fd, const T* __restrict__ psi,
                                                   const T* __restrict__ op, T* __restrict__ out) {
  const FoldDesc F = fd[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < F.Dout; i += (long long)gridDim.x * blockDim.x) {
    long long r = i;
    const int sp = (int)(r % F.d);
    r /= F.d;
    long long po = 0, oo = 0, ps = F.d, os = (long long)F.d * F.d;
    for (int a = 0; a < F.k; ++a) {
      const int nc = F.pchi[a] * F.achi[a];
      const int c = (int)(r % nc);
      r /= nc;
      po += (long long)(c % F.pchi[a]) * ps;
      oo += (long long)(c / F.pchi[a]) * os;
      ps *= F.pchi[a];
      os *= F.achi[a];
    }
    T acc = zerov<T>();
    for (int s = 0; s < F.d; ++s) acc += op[F.op_off + oo + sp + (long long)F.d * s] * psi[F.psi_off + po + s];
    out[F.out_off + i] = acc;
  }
}

}  // namespace

struct itnbp_lnet {
  itnbp_ctx* ctx = nullptr;
  int nv = 0, ne = 0;
  bool cplx = false;
  size_t es = 8;  // bytes per element
  std::vector<int> edges, inc_ptr, inc_edge, kchi, bchi, phys;
  std::vector<LVDesc> vdesc;
  std::vector<long long> msg_off;
  std::vector<int> msg_size;  // elements of the message pair of each edge
  long long msg_total = 0, ket_total = 0, bra_total = 0, maxD = 1;
  int maxP2 = 1;
  void* d_ket = nullptr;
  void* d_bra = nullptr;
  void* d_msg[2] = {nullptr, nullptr};
  int cur = 0;
  std::vector<char> msg_set;
  DevBuf d_vdesc, tasks, scratch, stats, small_out, d_moff, d_msize;
};

namespace {

itnbp_graph::Topo topo_of(const itnbp_lnet* n) {
  return itnbp_graph::Topo{n->nv, n->ne, n->edges.data(), n->inc_ptr.data(), n->inc_edge.data()};
}

// launch the tasks (scratch sized for them); msg_in/out may be the same buffer
// when no task reads a message another task of the launch writes
int run_layer_tasks(itnbp_lnet* L, const std::vector<LTask>& tasks, const void* msg_in, void* out, MsgStat* stats) {
  itnbp_ctx* ctx = L->ctx;
  if (tasks.empty()) return ITNBP_OK;
  const long long sstride = std::max<long long>(L->maxD, L->maxP2);
  const size_t n = tasks.size();
  ITNBP_RETCHK(ensure_buf(ctx, L->tasks, n * sizeof(LTask)));
  ITNBP_RETCHK(ensure_buf(ctx, L->scratch, n * 3 * (size_t)sstride * L->es));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(L->tasks.p, tasks.data(), n * sizeof(LTask), hipMemcpyHostToDevice, ctx->stream));
  if (L->cplx)
    hipLaunchKernelGGL(k_layer_task<cplx>, dim3((unsigned)n), dim3(LY_THREADS), 0, ctx->stream,
                       (const LVDesc*)L->d_vdesc.p, (const LTask*)L->tasks.p, (const cplx*)L->d_ket,
                       (const cplx*)L->d_bra, (const cplx*)msg_in, (cplx*)out, stats, (cplx*)L->scratch.p, sstride);
  else
    hipLaunchKernelGGL(k_layer_task<double>, dim3((unsigned)n), dim3(LY_THREADS), 0, ctx->stream,
                       (const LVDesc*)L->d_vdesc.p, (const LTask*)L->tasks.p, (const double*)L->d_ket,
                       (const double*)L->d_bra, (const double*)msg_in, (double*)out, stats,
                       (double*)L->scratch.p, sstride);
  ITNBP_HIPCHK(ctx, hipGetLastError());
  // the task table is reused by the next call: finish this launch first
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

int axis_into(const itnbp_lnet* L, int v, int e) {
  for (int a = L->inc_ptr[v]; a < L->inc_ptr[v + 1]; ++a)
    if (L->inc_edge[a] == e) return a - L->inc_ptr[v] + 1;
  return -1;
}

LTask message_task_l(const itnbp_lnet* L, int m, int slot, bool with_old, int normalize) {
  const itnbp_graph::Topo tp = topo_of(L);
  const int u = itnbp_graph::msg_src(tp, m);
  LTask t;
  t.v = u;
  t.j = axis_into(L, u, m / 2);
  t.slot = slot;
  t.normalize = normalize;
  t.out_off = L->msg_off[m];
  t.old_off = with_old ? L->msg_off[m] : -1;
  return t;
}

int lnet_update(itnbp_lnet* L, const itnbp_update_opts* o, itnbp_update_result* res) {
  itnbp_ctx* ctx = L->ctx;
  if (o->maxiter < 1) return itnbp_fail(ctx, ITNBP_ERR_ARG, "You need to specify a number of iterations for BP!");
  if (o->precision != ITNBP_PREC_F64)
    return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "the two-layer network runs in FP64 only");
  if (o->schedule != ITNBP_SCHED_SEQUENTIAL && o->schedule != ITNBP_SCHED_PARALLEL)
    return itnbp_fail(ctx, ITNBP_ERR_ARG, "unknown schedule");
  const itnbp_graph::Topo tp = topo_of(L);
  const bool seqsched = o->schedule == ITNBP_SCHED_SEQUENTIAL;
  std::vector<int> seq;
  if (o->nseq > 0) {
    if (!o->seq) return itnbp_fail(ctx, ITNBP_ERR_ARG, "nseq > 0 with a null seq");
    seq.assign(o->seq, o->seq + o->nseq);
  } else {
    itnbp_graph::default_sequence(tp, !seqsched, seq);
  }
  if (seq.empty()) {
    if (res) *res = itnbp_update_result{0, 0, 0.0, 0.0, 1, 0.0};
    return ITNBP_OK;
  }
  std::string why;
  const bool need_old = o->tol >= 0 || o->res_tol >= 0;
  if (const int rc = itnbp_graph::check_sequence(tp, L->msg_set, seq, seqsched, need_old, why))
    return itnbp_fail(ctx, rc, why);
  const size_t nseq = seq.size();
  ITNBP_RETCHK(ensure_buf(ctx, L->stats, nseq * sizeof(MsgStat)));
  const auto runs = seqsched ? itnbp_graph::sequential_runs(tp, seq)
                             : std::vector<std::pair<int, int>>{{0, (int)nseq}};
  std::vector<MsgStat> hs(nseq);
  itnbp_update_result r{};
  ITNBP_HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int it = 1; it <= o->maxiter; ++it) {
    void* cur = L->d_msg[L->cur];
    void* nxt = seqsched ? cur : L->d_msg[1 - L->cur];
    if (!seqsched)
      ITNBP_HIPCHK(ctx, hipMemcpyAsync(nxt, cur, L->msg_total * L->es, hipMemcpyDeviceToDevice, ctx->stream));
    for (const auto& run : runs) {
      std::vector<LTask> tasks;
      for (int i = run.first; i < run.second; ++i) {
        const int m = seq[i];
        tasks.push_back(message_task_l(L, m, i, L->msg_set[m] != 0, o->normalize));
      }
      ITNBP_RETCHK(run_layer_tasks(L, tasks, cur, nxt, (MsgStat*)L->stats.p));
      if (seqsched)
        for (int i = run.first; i < run.second; ++i) L->msg_set[seq[i]] = 1;
    }
    if (!seqsched) {
      L->cur = 1 - L->cur;
      for (int m : seq) L->msg_set[m] = 1;
    }
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(hs.data(), L->stats.p, nseq * sizeof(MsgStat), hipMemcpyDeviceToHost,
                                     ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    double sd = 0, mr = 0;
    for (size_t i = 0; i < nseq; ++i) {
      if (hs[i].bad > 0) return itnbp_fail(ctx, ITNBP_ERR_NONFINITE, "non-finite message");
      sd += hs[i].diff;
      mr = std::max(mr, hs[i].res2);
    }
    r.iterations = it;
    r.mean_diff = sd / (double)nseq;
    r.max_residual = std::sqrt(mr);
    if ((o->tol >= 0 && r.mean_diff <= o->tol) || (o->res_tol >= 0 && mr <= o->res_tol * o->res_tol)) {
      r.converged = 1;
      break;
    }
  }
  ITNBP_HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ITNBP_HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0;
  ITNBP_HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  r.device_ms = ms;
  ctx->path = "layered-f64" + std::string(L->cplx ? "-complex" : "");
  if (res) *res = r;
  return ITNBP_OK;
}

int lnet_region_scalars(itnbp_lnet* L, double* vert, double* edge) {
  itnbp_ctx* ctx = L->ctx;
  for (int v = 0; v < L->nv; ++v)
    for (int a = L->inc_ptr[v]; a < L->inc_ptr[v + 1]; ++a) {
      const int m = itnbp_graph::msg_id_into(topo_of(L), L->inc_edge[a], v);
      if (!L->msg_set[m])
        return itnbp_fail(ctx, ITNBP_ERR_STATE, "message " + std::to_string(m) + " is read before it is set");
    }
  const size_t nout = (size_t)L->nv + L->ne;
  ITNBP_RETCHK(ensure_buf(ctx, L->small_out, nout * L->es));
  std::vector<LTask> tasks(L->nv);
  for (int v = 0; v < L->nv; ++v) tasks[v] = LTask{v, -1, -1, 0, (long long)v, -1};
  const void* msg = L->d_msg[L->cur];
  ITNBP_RETCHK(run_layer_tasks(L, tasks, msg, L->small_out.p, nullptr));
  if (L->ne > 0) {
    char* eo = (char*)L->small_out.p + (size_t)L->nv * L->es;
    if (L->cplx)
      hipLaunchKernelGGL(k_layer_edge_scalar<cplx>, dim3(L->ne), dim3(64), 0, ctx->stream,
                         (const long long*)L->d_moff.p, (const int*)L->d_msize.p, (const cplx*)msg, (cplx*)eo);
    else
      hipLaunchKernelGGL(k_layer_edge_scalar<double>, dim3(L->ne), dim3(64), 0, ctx->stream,
                         (const long long*)L->d_moff.p, (const int*)L->d_msize.p, (const double*)msg, (double*)eo);
    ITNBP_HIPCHK(ctx, hipGetLastError());
  }
  std::vector<double> h(nout * (L->cplx ? 2 : 1));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(h.data(), L->small_out.p, nout * L->es, hipMemcpyDeviceToHost, ctx->stream));
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const int w = L->cplx ? 2 : 1;
  for (int v = 0; v < L->nv; ++v) {
    vert[2 * v] = h[(size_t)w * v];
    vert[2 * v + 1] = L->cplx ? h[(size_t)w * v + 1] : 0.0;
  }
  for (int e = 0; e < L->ne; ++e) {
    edge[2 * e] = h[(size_t)w * (L->nv + e)];
    edge[2 * e + 1] = L->cplx ? h[(size_t)w * (L->nv + e) + 1] : 0.0;
  }
  return ITNBP_OK;
}

}  // namespace

extern "C" {

int itnbp_lnet_create(itnbp_ctx* ctx, int nv, int ne, const int* edges, const int* inc_ptr, const int* inc_edge,
                      const int* ket_bond_dims, const int* bra_bond_dims, const int* phys_dims, int is_complex,
                      itnbp_lnet** out) {
  if (!ctx || !out) return ITNBP_ERR_ARG;
  *out = nullptr;
  return itnbp_guarded(ctx, [&]() -> int {
    const itnbp_graph::Topo tp{nv, ne, edges, inc_ptr, inc_edge};
    const itnbp_graph::Limits lim{ITNBP_MAXK, 64, 8, 1LL << 26};
    itnbp_graph::Layout lk, lb;
    std::string why;
    if (ne > 0 && !bra_bond_dims) return itnbp_fail(ctx, ITNBP_ERR_ARG, "null bra bond dimensions");
    if (const int rc = itnbp_graph::describe(tp, ket_bond_dims, phys_dims, lim, lk, why)) return itnbp_fail(ctx, rc, why);
    if (const int rc = itnbp_graph::describe(tp, bra_bond_dims, phys_dims, lim, lb, why)) return itnbp_fail(ctx, rc, why);
    auto L = std::make_unique<itnbp_lnet>();
    L->ctx = ctx;
    L->nv = nv;
    L->ne = ne;
    L->cplx = is_complex != 0;
    L->es = L->cplx ? 16 : 8;
    L->edges.assign(edges, edges + 2 * ne);
    L->inc_ptr.assign(inc_ptr, inc_ptr + nv + 1);
    L->inc_edge.assign(inc_edge, inc_edge + 2 * ne);
    L->kchi.assign(ket_bond_dims, ket_bond_dims + ne);
    L->bchi.assign(bra_bond_dims, bra_bond_dims + ne);
    L->phys.assign(phys_dims, phys_dims + nv);
    // message m of edge e: chi_ket(e) x chi_bra(e)
    L->msg_off.resize(2 * (size_t)ne);
    L->msg_size.resize(ne);
    for (int m = 0; m < 2 * ne; ++m) {
      L->msg_off[m] = L->msg_total;
      const int sz = L->kchi[m / 2] * L->bchi[m / 2];
      L->msg_total += sz;
      L->msg_size[m / 2] = sz;
      L->maxP2 = std::max(L->maxP2, sz);
    }
    L->vdesc.resize(nv);
    for (int v = 0; v < nv; ++v) {
      LVDesc& V = L->vdesc[v];
      V = LVDesc{};
      V.k = inc_ptr[v + 1] - inc_ptr[v];
      V.kd[0] = V.bd[0] = phys_dims[v];
      long long Dmax = phys_dims[v];
      for (int a = 1; a <= V.k; ++a) {
        const int e = inc_edge[inc_ptr[v] + a - 1];
        V.kd[a] = L->kchi[e];
        V.bd[a] = L->bchi[e];
        V.in_msg[a] = L->msg_off[itnbp_graph::msg_id_into(tp, e, v)];
        Dmax *= std::max(V.kd[a], V.bd[a]);
      }
      V.ket_off = lk.psi_off[v];
      V.bra_off = lb.psi_off[v];
      V.Dk = lk.site_D[v];
      L->maxD = std::max(L->maxD, Dmax);
      L->maxP2 = std::max<int>(L->maxP2, phys_dims[v] * phys_dims[v]);
    }
    L->ket_total = lk.psi_total;
    L->bra_total = lb.psi_total;
    L->msg_set.assign(2 * ne, 0);
    ITNBP_HIPCHK(ctx, hipSetDevice(ctx->device));
    if (hipMalloc(&L->d_ket, std::max<size_t>(16, L->ket_total * L->es)) != hipSuccess ||
        hipMalloc(&L->d_bra, std::max<size_t>(16, L->bra_total * L->es)) != hipSuccess ||
        hipMalloc(&L->d_msg[0], std::max<size_t>(16, L->msg_total * L->es)) != hipSuccess ||
        hipMalloc(&L->d_msg[1], std::max<size_t>(16, L->msg_total * L->es)) != hipSuccess) {
      (void)hipGetLastError();
      itnbp_lnet_destroy(L.release());
      return itnbp_fail(ctx, ITNBP_ERR_NOMEM, "device allocation for the two-layer network failed");
    }
    ITNBP_RETCHK(upload(ctx, L->d_vdesc, L->vdesc.data(), sizeof(LVDesc) * nv));
    if (ne > 0) {
      ITNBP_RETCHK(upload(ctx, L->d_moff, L->msg_off.data(), 8 * L->msg_off.size()));
      ITNBP_RETCHK(upload(ctx, L->d_msize, L->msg_size.data(), 4 * L->msg_size.size()));
    }
    ITNBP_HIPCHK(ctx, hipMemsetAsync(L->d_ket, 0, L->ket_total * L->es, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemsetAsync(L->d_bra, 0, L->bra_total * L->es, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemsetAsync(L->d_msg[0], 0, L->msg_total * L->es, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *out = L.release();
    return ITNBP_OK;
  });
}

int itnbp_lnet_destroy(itnbp_lnet* L) {
  if (!L) return ITNBP_ERR_ARG;
  if (L->ctx) (void)hipStreamSynchronize(L->ctx->stream);
  for (void* p : {L->d_ket, L->d_bra, L->d_msg[0], L->d_msg[1]})
    if (p) (void)hipFree(p);
  for (DevBuf* b : {&L->d_vdesc, &L->tasks, &L->scratch, &L->stats, &L->small_out, &L->d_moff, &L->d_msize})
    free_buf(*b);
  delete L;
  return ITNBP_OK;
}

int64_t itnbp_lnet_msg_numel(itnbp_lnet* L) { return L ? L->msg_total : -1; }
int64_t itnbp_lnet_ket_numel(itnbp_lnet* L) { return L ? L->ket_total : -1; }
int64_t itnbp_lnet_bra_numel(itnbp_lnet* L) { return L ? L->bra_total : -1; }

int itnbp_lnet_set_tensors(itnbp_lnet* L, const double* ket, const double* bra) {
  if (!L || !ket || !bra) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(L->d_ket, ket, L->ket_total * L->es, hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(L->d_bra, bra, L->bra_total * L->es, hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

int itnbp_lnet_fold_operator(itnbp_lnet* L, const double* psi, const int* psi_bond_dims, const double* ops,
                             const int* op_bond_dims) {
  if (!L || !psi || !ops || (L->ne > 0 && (!psi_bond_dims || !op_bond_dims))) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    for (int e = 0; e < L->ne; ++e)
      if (psi_bond_dims[e] < 1 || op_bond_dims[e] < 1 || psi_bond_dims[e] * op_bond_dims[e] != L->kchi[e])
        return itnbp_fail(ctx, ITNBP_ERR_ARG, "ket bond dimension must be chi_psi * chi_op on every edge");
    std::vector<FoldDesc> fd(L->nv);
    long long po = 0, oo = 0;
    for (int v = 0; v < L->nv; ++v) {
      FoldDesc& F = fd[v];
      F = FoldDesc{};
      F.k = L->inc_ptr[v + 1] - L->inc_ptr[v];
      F.d = L->phys[v];
      long long Dp = F.d, Do = (long long)F.d * F.d;
      for (int a = 0; a < F.k; ++a) {
        const int e = L->inc_edge[L->inc_ptr[v] + a];
        F.pchi[a] = psi_bond_dims[e];
        F.achi[a] = op_bond_dims[e];
        Dp *= F.pchi[a];
        Do *= F.achi[a];
      }
      F.psi_off = po;
      F.op_off = oo;
      F.out_off = L->vdesc[v].ket_off;
      F.Dout = L->vdesc[v].Dk;
      po += Dp;
      oo += Do;
    }
    DevBuf dpsi, dop, dfd;
    struct G {
      DevBuf* b[3];
      ~G() {
        for (DevBuf* x : b) free_buf(*x);
      }
    } guard{{&dpsi, &dop, &dfd}};
    ITNBP_RETCHK(upload(ctx, dpsi, psi, po * L->es));
    ITNBP_RETCHK(upload(ctx, dop, ops, oo * L->es));
    ITNBP_RETCHK(upload(ctx, dfd, fd.data(), fd.size() * sizeof(FoldDesc)));
    if (L->cplx)
      hipLaunchKernelGGL(k_layer_fold<cplx>, dim3(16, L->nv), dim3(256), 0, ctx->stream, (const FoldDesc*)dfd.p,
                         (const cplx*)dpsi.p, (const cplx*)dop.p, (cplx*)L->d_ket);
    else
      hipLaunchKernelGGL(k_layer_fold<double>, dim3(16, L->nv), dim3(256), 0, ctx->stream, (const FoldDesc*)dfd.p,
                         (const double*)dpsi.p, (const double*)dop.p, (double*)L->d_ket);
    ITNBP_HIPCHK(ctx, hipGetLastError());
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

int itnbp_lnet_get_ket(itnbp_lnet* L, double* ket) {
  if (!L || !ket) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(ket, L->d_ket, L->ket_total * L->es, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

int itnbp_lnet_set_message(itnbp_lnet* L, int m, const double* data) {
  if (!L || !data || m < 0 || m >= 2 * L->ne) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    const size_t bytes = (size_t)L->msg_size[m / 2] * L->es;
    ITNBP_HIPCHK(ctx, hipMemcpyAsync((char*)L->d_msg[L->cur] + L->msg_off[m] * L->es, data, bytes,
                                     hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    L->msg_set[m] = 1;
    return ITNBP_OK;
  });
}

int itnbp_lnet_get_message(itnbp_lnet* L, int m, double* data) {
  if (!L || !data || m < 0 || m >= 2 * L->ne) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    if (!L->msg_set[m])
      return itnbp_fail(ctx, ITNBP_ERR_STATE, "message " + std::to_string(m) + " has not been set");
    const size_t bytes = (size_t)L->msg_size[m / 2] * L->es;
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(data, (char*)L->d_msg[L->cur] + L->msg_off[m] * L->es, bytes,
                                     hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

int itnbp_lnet_message_is_set(itnbp_lnet* L, int m) {
  if (!L || m < 0 || m >= 2 * L->ne) return -1;
  return L->msg_set[m] ? 1 : 0;
}

int itnbp_lnet_init_identity_messages(itnbp_lnet* L) {
  if (!L) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    for (int e = 0; e < L->ne; ++e)
      if (L->kchi[e] != L->bchi[e])
        return itnbp_fail(ctx, ITNBP_ERR_ARG, "identity messages need equal ket and bra bond dimensions");
    const int w = L->cplx ? 2 : 1;
    std::vector<double> h((size_t)L->msg_total * w, 0.0);
    for (int m = 0; m < 2 * L->ne; ++m) {
      const int n = L->kchi[m / 2];
      for (int i = 0; i < n; ++i) h[(size_t)w * (L->msg_off[m] + i + (long long)n * i)] = 1.0;
    }
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(L->d_msg[L->cur], h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    std::fill(L->msg_set.begin(), L->msg_set.end(), 1);
    return ITNBP_OK;
  });
}

int itnbp_lnet_update(itnbp_lnet* L, const itnbp_update_opts* o, itnbp_update_result* res) {
  if (!L || !o) return ITNBP_ERR_ARG;
  return itnbp_guarded(L->ctx, [&]() -> int { return lnet_update(L, o, res); });
}

int itnbp_lnet_region_scalars(itnbp_lnet* L, double* vert, double* edge) {
  if (!L || !vert || (L->ne > 0 && !edge)) return ITNBP_ERR_ARG;
  return itnbp_guarded(L->ctx, [&]() -> int { return lnet_region_scalars(L, vert, edge); });
}

int itnbp_lnet_logscalar(itnbp_lnet* L, double* re, double* im) {
  if (!L || !re || !im) return ITNBP_ERR_ARG;
  return itnbp_guarded(L->ctx, [&]() -> int {
    std::vector<double> vs(2 * (size_t)L->nv), es(2 * (size_t)std::max(1, L->ne));
    ITNBP_RETCHK(lnet_region_scalars(L, vs.data(), es.data()));
    // sum(log.(vertex scalars)) - sum(log.(edge scalars)) (abstractbeliefpropagationcache.jl:397-412),
    // complex logarithms
    double sr = 0, si = 0;
    auto acc = [&](double x, double y, double sign) {
      sr += sign * 0.5 * std::log(x * x + y * y);
      si += sign * std::atan2(y, x);
    };
    for (int v = 0; v < L->nv; ++v) acc(vs[2 * v], vs[2 * v + 1], 1.0);
    for (int e = 0; e < L->ne; ++e) acc(es[2 * e], es[2 * e + 1], -1.0);
    *re = sr;
    *im = si;
    return ITNBP_OK;
  });
}

int itnbp_lnet_rdm1(itnbp_lnet* L, int n, const int* verts, double* out) {
  if (!L || n < 0 || (n > 0 && (!verts || !out))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  itnbp_ctx* ctx = L->ctx;
  return itnbp_guarded(ctx, [&]() -> int {
    std::vector<LTask> tasks(n);
    std::vector<long long> offs(n);
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
      const int v = verts[i];
      if (v < 0 || v >= L->nv) return itnbp_fail(ctx, ITNBP_ERR_ARG, "vertex out of range");
      for (int a = L->inc_ptr[v]; a < L->inc_ptr[v + 1]; ++a)
        if (!L->msg_set[itnbp_graph::msg_id_into(topo_of(L), L->inc_edge[a], v)])
          return itnbp_fail(ctx, ITNBP_ERR_STATE, "an incoming message of the vertex is not set");
      offs[i] = tot;
      tasks[i] = LTask{v, 0, -1, 0, tot, -1};
      tot += (long long)L->phys[v] * L->phys[v];
    }
    ITNBP_RETCHK(ensure_buf(ctx, L->small_out, tot * L->es));
    ITNBP_RETCHK(run_layer_tasks(L, tasks, L->d_msg[L->cur], L->small_out.p, nullptr));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(out, L->small_out.p, tot * L->es, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

}  // extern "C"
