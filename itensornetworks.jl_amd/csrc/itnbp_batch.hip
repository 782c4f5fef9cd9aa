// Batched heterogeneous BP: many independent networks (any lattices, bond
// dimensions and degrees, one storage dtype) updated together.  The instances
// are merged on the device into one disjoint-union network, so every message of
// every instance is a task of the same per-size-class launches of a sweep;
// each instance's convergence is tested on its own statistics and converged
// instances drop out of the launch plan.  Semantics per instance are those of
// update(bpc; alg="bp") with the parallel schedule
// (src/caches/abstractbeliefpropagationcache.jl:294-337, src/edge_sequences.jl:49-51).
#include <algorithm>
#include <cmath>
#include <vector>

#include "itnbp_ops.hpp"

namespace {

// one instance of the merged network: its statistics slots and message range
struct BatchInst {
  int slot0, nslot;
  long long moff, mlen;
};

// per-instance result, written on the device by k_batch_converge
struct BatchOut {
  int iters, conv, bad, pad;
  double mean_diff, max_res2;
};

// Carries the messages of finished instances from the current buffer to the
// next (the per-vertex kernels skip them); one workgroup per instance.
template <typename TM>
__global__ __launch_bounds__(256) void k_batch_carry(const BatchInst* __restrict__ inst, const int* __restrict__ active,
                                                     const TM* __restrict__ cur, TM* __restrict__ nxt) {
  const int i = blockIdx.x;
  if (active[i]) return;
  const long long o = inst[i].moff, n = inst[i].mlen;
  for (long long j = threadIdx.x; j < n; j += blockDim.x) nxt[o + j] = cur[o + j];
}

// The convergence test of update(bpc) per instance after a sweep
// (src/caches/abstractbeliefpropagationcache.jl:323-333: the mean message_diff
// against tol, here also the largest residual against res_tol), on the device:
// one wave per instance; a finished instance (converged, non-finite, or at
// maxiter) is switched off for the following sweeps.
__global__ __launch_bounds__(64) void k_batch_converge(const MsgStat* __restrict__ stats,
                                                      const BatchInst* __restrict__ inst, int* __restrict__ active,
                                                      BatchOut* __restrict__ out, const long long* __restrict__ ctr,
                                                      int maxiter, double tol, double res_tol) {
  const int i = blockIdx.x, l = threadIdx.x;
  if (!active[i]) return;
  const int s0 = inst[i].slot0, ns = inst[i].nslot;
  double sd = 0, mr = 0;
  int bad = 0;
  for (int m = l; m < ns; m += 64) {
    const MsgStat st = stats[s0 + m];
    sd += st.diff;
    mr = fmax(mr, st.res2);
    bad |= st.bad > 0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sd += __shfl_xor(sd, o);
    mr = fmax(mr, __shfl_xor(mr, o));
    bad |= __shfl_xor(bad, o);
  }
  if (l == 0) {
    const int it = (int)(*ctr) + 1;
    const double mean = ns ? sd / ns : 0.0;
    BatchOut r = out[i];
    r.iters = it;
    r.mean_diff = mean;
    r.max_res2 = mr;
    r.bad = bad;
    const bool stop = (tol >= 0 && mean <= tol) || (res_tol >= 0 && mr <= res_tol * res_tol);
    r.conv = stop ? 1 : 0;
    out[i] = r;
    if (stop || bad || it >= maxiter) active[i] = 0;
  }
}

__global__ void k_batch_count(long long* ctr) {
  if (threadIdx.x == 0) *ctr += 1;
}

// Device-resident sweep loop for the batched update when every message runs on
// the per-vertex kernels: K Jacobi sweeps (carry, vertex kernels, convergence,
// sweep count) are captured once into a HIP graph and replayed until every
// instance has finished; the host reads the active flags and results back once
// per replay.  The first sweep runs eagerly (it also sets the kernels' launch
// attributes, which capture must not do).
constexpr int BATCH_GRAPH_SWEEPS = 8;  // even: the buffers return to their roles

int batched_graph_loop(itnbp_net* M, int n, const itnbp_update_opts* o, int acc_f64, void* const buf[2],
                       const BatchInst* d_inst, int* d_active, BatchOut* d_out, long long* d_ctr,
                       std::vector<BatchOut>& hout, int& sweeps_run, bool& graphed) {
  itnbp_ctx* ctx = M->ctx;
  MsgStat* d_stats = (MsgStat*)M->stats.p;
  auto sweep = [&](int c) -> hipError_t {
    const void* cur = buf[c];
    void* nxt = buf[1 - c];
    if (acc_f64)
      hipLaunchKernelGGL(k_batch_carry<double>, dim3(n), dim3(256), 0, ctx->stream, d_inst, d_active,
                         (const double*)cur, (double*)nxt);
    else
      hipLaunchKernelGGL(k_batch_carry<float>, dim3(n), dim3(256), 0, ctx->stream, d_inst, d_active,
                         (const float*)cur, (float*)nxt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (small_sweep(M, acc_f64, M->d_psi, M->dtype, cur, nxt, d_stats, o->normalize) != ITNBP_OK)
      return hipErrorLaunchFailure;
    hipLaunchKernelGGL(k_batch_converge, dim3(n), dim3(64), 0, ctx->stream, d_stats, d_inst, d_active, d_out,
                       (const long long*)d_ctr, o->maxiter, o->tol, o->res_tol);
    hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(64), 0, ctx->stream, d_ctr);
    return hipGetLastError();
  };
  std::vector<int> hact(n);
  auto readback = [&]() -> int {
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(hact.data(), d_active, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(hout.data(), d_out, n * sizeof(BatchOut), hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i)
      if (hout[i].bad)
        return itnbp_fail(ctx, ITNBP_ERR_NONFINITE, "non-finite message in batched instance " + std::to_string(i));
    return ITNBP_OK;
  };
  auto any_active = [&]() { return std::any_of(hact.begin(), hact.end(), [](int a) { return a != 0; }); };
  hipError_t e = sweep(M->cur);
  if (e != hipSuccess) return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("batched sweep: ") + hipGetErrorString(e));
  M->cur = 1 - M->cur;
  sweeps_run = 1;
  graphed = false;
  ITNBP_RETCHK(readback());
  if (!any_active()) return ITNBP_OK;
  // the profiler's event records stay out of the graph
  const bool prof = ctx->prof;
  ctx->prof = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t ec = hipSuccess;
    for (int k = 0; k < BATCH_GRAPH_SWEEPS && ec == hipSuccess; ++k) ec = sweep((M->cur + k) & 1);
    e = hipStreamEndCapture(ctx->stream, &graph);
    if (ec != hipSuccess) e = ec;
  }
  ctx->prof = prof;
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  int rc = ITNBP_OK;
  if (e != hipSuccess) {
    // no graph (capture refused): the same sweeps launched one by one
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGetLastError();
    graphed = false;
    while (rc == ITNBP_OK && sweeps_run < o->maxiter && any_active()) {
      e = sweep(M->cur);
      if (e != hipSuccess) return itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("batched sweep: ") + hipGetErrorString(e));
      M->cur = 1 - M->cur;
      ++sweeps_run;
      rc = readback();
    }
    return rc;
  }
  graphed = true;
  while (sweeps_run < o->maxiter && any_active()) {
    e = hipGraphLaunch(exec, ctx->stream);
    if (e != hipSuccess) {
      rc = itnbp_fail(ctx, ITNBP_ERR_HIP, std::string("batched graph launch: ") + hipGetErrorString(e));
      break;
    }
    sweeps_run += BATCH_GRAPH_SWEEPS;
    if ((rc = readback()) != ITNBP_OK) break;
  }
  (void)hipGraphExecDestroy(exec);
  return rc;
}

bool batched_graph_enabled() {
  const char* e = getenv("ITNBP_BATCH_GRAPH");
  return !(e && e[0] == '0');
}

int update_batched(int n, itnbp_net** nets, const itnbp_update_opts* o, itnbp_update_result* res) {
  itnbp_ctx* ctx = nets[0]->ctx;
  if (o->maxiter < 1) return itnbp_fail(ctx, ITNBP_ERR_ARG, "You need to specify a number of iterations for BP!");
  if (o->schedule != ITNBP_SCHED_PARALLEL || o->nseq != 0)
    return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED, "batched update runs the default parallel schedule");
  // FP64, or FP32 arithmetic for ITNBP_PREC_F32, ITNBP_PREC_BF16X3 and
  // ITNBP_PREC_F16X2 (FP32-class results; there are no 16-bit batched kernels);
  // every phased or refining mode is refused
  if (o->precision != ITNBP_PREC_F64 && o->precision != ITNBP_PREC_F32 && o->precision != ITNBP_PREC_BF16X3 &&
      o->precision != ITNBP_PREC_F16X2)
    return itnbp_fail(ctx, ITNBP_ERR_UNSUPPORTED,
                      "batched update supports precision f64, and f32, bf16x3 and f16x2 (FP32 arithmetic) only");
  const int dtype = nets[0]->dtype;
  // merged description
  std::vector<int> edges, inc_ptr{0}, inc_edge, chi, phys;
  std::vector<int> vbase(n), ebase(n);
  int NV = 0, NE = 0;
  for (int i = 0; i < n; ++i) {
    itnbp_net* t = nets[i];
    if (!t || t->ctx != ctx || t->dtype != dtype)
      return itnbp_fail(ctx, ITNBP_ERR_ARG, "batched networks must share the context and the dtype");
    for (int m = 0; m < 2 * t->ne; ++m)
      if (!t->msg_set[m]) return itnbp_fail(ctx, ITNBP_ERR_STATE, "every message of a batched network must be set");
    vbase[i] = NV;
    ebase[i] = NE;
    for (int e = 0; e < t->ne; ++e) {
      edges.push_back(t->edges[2 * e] + NV);
      edges.push_back(t->edges[2 * e + 1] + NV);
      chi.push_back(t->chi[e]);
    }
    for (int v = 0; v < t->nv; ++v) {
      for (int a = t->inc_ptr[v]; a < t->inc_ptr[v + 1]; ++a) inc_edge.push_back(t->inc_edge[a] + NE);
      inc_ptr.push_back((int)inc_edge.size());
      phys.push_back(t->phys[v]);
    }
    NV += t->nv;
    NE += t->ne;
  }
  itnbp_net* M = nullptr;
  ITNBP_RETCHK(itnbp_net_create(ctx, NV, NE, edges.data(), inc_ptr.data(), inc_edge.data(), chi.data(),
                                phys.data(), dtype, &M));
  std::unique_ptr<itnbp_net, int (*)(itnbp_net*)> guard(M, itnbp_net_destroy);
  // copy site tensors and messages in (instance blocks are contiguous in the merged layout)
  const size_t eb = dtype_size(dtype);
  std::vector<long long> pbase(n), mbase(n);
  for (int i = 0; i < n; ++i) {
    itnbp_net* t = nets[i];
    pbase[i] = M->psi_off[vbase[i]];
    mbase[i] = M->msg_off[2 * ebase[i]];
    ITNBP_HIPCHK(ctx, hipMemcpyAsync((char*)M->d_psi + pbase[i] * eb, t->d_psi, t->psi_total * eb,
                                     hipMemcpyDeviceToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(M->d_msg[0] + mbase[i], t->d_msg[t->cur], t->msg_total * 8,
                                     hipMemcpyDeviceToDevice, ctx->stream));
  }
  std::fill(M->msg_set.begin(), M->msg_set.end(), 1);
  M->cur = 0;
  // the per-vertex eigenbasis kernel needs symmetric messages in every instance
  M->msgs_sym = std::all_of(nets, nets + n, [](const itnbp_net* t) { return t->msgs_sym; });
  const int acc_f64 = o->precision == ITNBP_PREC_F64 ? 1 : 0;
  void* buf[2];
  if (acc_f64) {
    buf[0] = M->d_msg[0];
    buf[1] = M->d_msg[1];
  } else {
    for (int b = 0; b < 2; ++b)
      ITNBP_HIPCHK(ctx, hipMalloc((void**)&M->d_msg32[b], std::max<size_t>(16, M->msg_total * 4)));
    ITNBP_HIPCHK(ctx, launch_convert_msgs(M->d_msg[0], M->d_msg32[0], M->msg_total, ctx->stream));
    buf[0] = M->d_msg32[0];
    buf[1] = M->d_msg32[1];
  }
  const size_t mb = acc_f64 ? 8 : 4;
  // tasks: slot = merged message id; instance i owns slots [2 ebase[i], 2 ebase[i] + 2 ne_i)
  std::vector<char> active(n, 1);
  std::vector<itnbp_update_result> out(n);
  for (int i = 0; i < n; ++i) out[i] = itnbp_update_result{0, 0, 0.0, 0.0, 0, 0.0};
  ITNBP_RETCHK(ensure_buf(ctx, M->stats, sizeof(MsgStat) * (size_t)std::max(1, 2 * NE)));
  MsgStat* d_stats = (MsgStat*)M->stats.p;
  std::vector<MsgStat> hs(2 * NE);
  TaskPlan plan;
  bool replan = true, use_small = false;
  int small_first = -1;  // vertex tasks of the first plan (reported in the path)
  bool graph_mode = false, graphed = false;
  int tile = 0;
  // device-resident loop: every message of every instance on the per-vertex
  // kernels (chi <= 16, symmetric messages)
  DevBuf d_inst, d_active, d_out, d_ctr;
  struct BufGuard {
    DevBuf* b[4];
    ~BufGuard() {
      for (DevBuf* x : b) free_buf(*x);
    }
  } bguard{{&d_inst, &d_active, &d_out, &d_ctr}};
  if (batched_graph_enabled() && M->msgs_sym) {
    std::vector<int> vinst(NV);
    for (int i = 0; i < n; ++i)
      std::fill(vinst.begin() + vbase[i], vinst.begin() + vbase[i] + nets[i]->nv, i);
    ITNBP_RETCHK(ensure_buf(ctx, d_inst, n * sizeof(BatchInst)));
    ITNBP_RETCHK(ensure_buf(ctx, d_active, n * sizeof(int)));
    ITNBP_RETCHK(ensure_buf(ctx, d_out, n * sizeof(BatchOut)));
    ITNBP_RETCHK(ensure_buf(ctx, d_ctr, sizeof(long long)));
    small_set_batch(M, vinst, (const int*)d_active.p, (const long long*)d_ctr.p);
    std::vector<int> seq(2 * NE);
    for (int m = 0; m < 2 * NE; ++m) seq[m] = m;
    std::vector<char> handled(2 * NE, 0);
    ITNBP_RETCHK(small_prepare(M, seq, acc_f64, handled));
    graph_mode = std::all_of(handled.begin(), handled.end(), [](char h) { return h != 0; });
    if (!graph_mode) small_set_batch(M, {}, nullptr, nullptr);
  }
  if (graph_mode) {
    std::vector<BatchInst> hinst(n);
    for (int i = 0; i < n; ++i)
      hinst[i] = BatchInst{2 * ebase[i], 2 * nets[i]->ne, mbase[i], nets[i]->msg_total};
    std::vector<int> ones(n, 1);
    std::vector<BatchOut> hout(n, BatchOut{0, 0, 0, 0, 0.0, 0.0});
    const long long zero = 0;
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_inst.p, hinst.data(), n * sizeof(BatchInst), hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_active.p, ones.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_out.p, hout.data(), n * sizeof(BatchOut), hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(d_ctr.p, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
    ITNBP_RETCHK(small_autotune(M, acc_f64, M->d_psi, dtype, buf[M->cur], o->normalize));
    tile = small_big_tile(M);
    small_first = small_task_count(M);
    int sweeps = 0;
    ITNBP_HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    ITNBP_RETCHK(batched_graph_loop(M, n, o, acc_f64, buf, (const BatchInst*)d_inst.p, (int*)d_active.p,
                                    (BatchOut*)d_out.p, (long long*)d_ctr.p, hout, sweeps, graphed));
    for (int i = 0; i < n; ++i) {
      out[i].iterations = hout[i].iters;
      out[i].mean_diff = hout[i].mean_diff;
      out[i].max_residual = std::sqrt(hout[i].max_res2);
      out[i].converged = hout[i].conv;
    }
  }
  if (!graph_mode) ITNBP_HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  for (int it = 1; !graph_mode && it <= o->maxiter; ++it) {
    if (replan) {
      // whole-vertex tasks (every chi <= 16, site in LDS) of the active instances on
      // the per-vertex kernel, every other message on the generic plan; the
      // statistics slot of a message is its merged id
      std::vector<int> seq(2 * NE);
      for (int m = 0; m < 2 * NE; ++m) seq[m] = m;
      std::vector<char> handled(2 * NE, 0);
      for (int i = 0; i < n; ++i)
        if (!active[i])
          std::fill(handled.begin() + 2 * ebase[i], handled.begin() + 2 * (ebase[i] + nets[i]->ne), 1);
      ITNBP_RETCHK(small_prepare(M, seq, acc_f64, handled));
      use_small = small_task_count(M) > 0;
      if (small_first < 0) small_first = small_task_count(M);
      std::vector<OcTask> tasks;
      for (int i = 0; i < n; ++i) {
        if (!active[i]) continue;
        for (int m = 2 * ebase[i]; m < 2 * (ebase[i] + nets[i]->ne); ++m)
          if (!handled[m]) tasks.push_back(message_task(M, m, m, true, o->normalize));
      }
      plan = TaskPlan();
      ITNBP_RETCHK(plan_tasks(M, tasks, acc_f64, plan));
      ITNBP_RETCHK(upload_plan(M, plan, M->tasks));
      replan = false;
    }
    void* cur = buf[M->cur];
    void* nxt = buf[1 - M->cur];
    // inactive instances keep their messages: carry the whole buffer over first
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(nxt, cur, M->msg_total * mb, hipMemcpyDeviceToDevice, ctx->stream));
    if (use_small) ITNBP_RETCHK(small_sweep(M, acc_f64, M->d_psi, dtype, cur, nxt, d_stats, o->normalize));
    ITNBP_RETCHK(launch_plan(M, plan, 0, plan.groups.size(), (const OcTask*)M->tasks.p, M->d_psi, dtype, cur, nxt,
                             d_stats));
    M->cur = 1 - M->cur;
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(hs.data(), d_stats, hs.size() * sizeof(MsgStat), hipMemcpyDeviceToHost,
                                     ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    bool any = false;
    for (int i = 0; i < n; ++i) {
      if (!active[i]) continue;
      const int ne2 = 2 * nets[i]->ne;
      double sd = 0, mr = 0;
      bool bad = false;
      for (int m = 2 * ebase[i]; m < 2 * ebase[i] + ne2; ++m) {
        sd += hs[m].diff;
        mr = std::max(mr, hs[m].res2);
        bad = bad || hs[m].bad > 0;
      }
      if (bad) return itnbp_fail(ctx, ITNBP_ERR_NONFINITE, "non-finite message in batched instance " + std::to_string(i));
      out[i].iterations = it;
      out[i].mean_diff = ne2 ? sd / ne2 : 0.0;
      out[i].max_residual = std::sqrt(mr);
      bool stop = false;
      if (o->tol >= 0 && out[i].mean_diff <= o->tol) stop = true;
      if (o->res_tol >= 0 && mr <= o->res_tol * o->res_tol) stop = true;
      if (stop) {
        out[i].converged = 1;
        active[i] = 0;
        replan = true;
      }
      any = any || active[i];
    }
    if (!any) break;
  }
  ITNBP_HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  ITNBP_HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0;
  ITNBP_HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  // scatter the messages back (FP64 canonical buffers)
  if (!acc_f64)
    ITNBP_HIPCHK(ctx, launch_convert_msgs_back(M->d_msg32[M->cur], M->d_msg[M->cur], M->msg_total, ctx->stream));
  for (int i = 0; i < n; ++i) {
    itnbp_net* t = nets[i];
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(t->d_msg[t->cur], M->d_msg[M->cur] + mbase[i], t->msg_total * 8,
                                     hipMemcpyDeviceToDevice, ctx->stream));
    out[i].device_ms = ms;
    if (res) res[i] = out[i];
  }
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (graph_mode)
    ctx->path = std::string(graphed ? "batched-graph" : "batched-device") + " [vertex " + std::to_string(small_first) +
                (tile ? std::string(", ") + small_big_kind(M) + " tile " + std::to_string(tile) : "") +
                (small_queue_used(M) ? ", persistent" : "") + "]";
  else
    ctx->path = small_first > 0 ? "batched-generic [vertex " + std::to_string(small_first) + "]" : "batched-generic";
  return ITNBP_OK;
}

}  // namespace

extern "C" int itnbp_update_batched(int n, itnbp_net** nets, const itnbp_update_opts* o,
                                    itnbp_update_result* res) {
  if (n <= 0 || !nets || !o || !nets[0]) return ITNBP_ERR_ARG;
  return itnbp_guarded(nets[0]->ctx, [&]() -> int { return update_batched(n, nets, o, res); });
}
