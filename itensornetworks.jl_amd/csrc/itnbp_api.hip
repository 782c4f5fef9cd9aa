// C ABI of libitnbp: contexts, networks, messages, the BP sweep driver,
// environments / expectation values / scalars, normalisation.
//
// Reference semantics (paths relative to the ITensorNetworks.jl repository):
//   update(bpc; alg="bp", ...)       src/caches/abstractbeliefpropagationcache.jl:272-337
//   updated_message("contract")      src/caches/abstractbeliefpropagationcache.jl:225-239
//   identity_messages                src/formnetworks/quadraticformnetwork.jl:96-124
//   default edge sequence            src/edge_sequences.jl:9-51
//   environment / region_scalar      src/caches/beliefpropagationcache.jl:100-119
//   expect                           src/expect.jl:5-41
//   logscalar / rescale              src/caches/abstractbeliefpropagationcache.jl:339-408,
//                                    src/caches/beliefpropagationcache.jl:121-139
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "itnbp_graph.hpp"
#include "itnbp_internal.hpp"
#include "itnbp_ops.hpp"

namespace {

int set_err(itnbp_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                               \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      return set_err((ctx), ITNBP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define RETCHK(call)          \
  do {                        \
    int _r = (call);          \
    if (_r != ITNBP_OK) return _r; \
  } while (0)

template <class F>
int guarded(itnbp_ctx* ctx, F&& f) {
  return itnbp_guarded(ctx, static_cast<F&&>(f));
}

}  // namespace

// ---- per-kernel timing ----------------------------------------------------------
int prof_mark(itnbp_ctx* ctx) {
  if (!ctx || !ctx->prof) return -1;
  if (ctx->ev_used == ctx->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    ctx->ev_pool.push_back(e);
  }
  const int i = (int)ctx->ev_used++;
  if (hipEventRecord(ctx->ev_pool[i], ctx->stream) != hipSuccess) return -1;
  return i;
}

void prof_add(itnbp_ctx* ctx, const char* name, int e0, int e1) {
  if (!ctx || !ctx->prof || e0 < 0 || e1 < 0) return;
  ctx->recs.push_back(ProfRec{name, e0, e1});
}

// ---- shared helpers (declared in itnbp_ops.hpp) ------------------------------
size_t dtype_size(int dtype) { return dtype == ITNBP_F64 ? 8 : dtype == ITNBP_F32 ? 4 : 2; }

int itnbp_fail(itnbp_ctx* ctx, int code, const std::string& msg) { return set_err(ctx, code, msg); }

int ensure_buf(itnbp_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.bytes >= bytes) return ITNBP_OK;
  if (b.p) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(b.p));
  }
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(ctx, ITNBP_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  b.bytes = bytes;
  return ITNBP_OK;
}

void free_buf(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

int upload(itnbp_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
  RETCHK(ensure_buf(ctx, b, bytes));
  HIPCHK(ctx, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return ITNBP_OK;
}

itnbp_graph::Topo topo_of(const itnbp_net* net) {
  return itnbp_graph::Topo{net->nv, net->ne, net->edges.data(), net->inc_ptr.data(), net->inc_edge.data()};
}
int msg_id_into(const itnbp_net* net, int e, int to) { return itnbp_graph::msg_id_into(topo_of(net), e, to); }
int msg_src(const itnbp_net* net, int m) { return itnbp_graph::msg_src(topo_of(net), m); }
int msg_dst(const itnbp_net* net, int m) { return itnbp_graph::msg_dst(topo_of(net), m); }

int axis_of(const itnbp_net* net, int v, int e) {
  for (int a = net->inc_ptr[v]; a < net->inc_ptr[v + 1]; ++a)
    if (net->inc_edge[a] == e) return a - net->inc_ptr[v] + 1;
  return -1;
}

int all_bonds_mask(const itnbp_net* net, int v) {
  int k = net->inc_ptr[v + 1] - net->inc_ptr[v];
  return ((1 << (k + 1)) - 1) & ~1;
}

OcTask message_task(const itnbp_net* net, int m, int slot, bool with_old, int normalize) {
  OcTask t;
  const int u = msg_src(net, m);
  t.v = u;
  t.j = axis_of(net, u, m / 2);
  t.s_open = 0;
  t.absorb_mask = all_bonds_mask(net, u) & ~(1 << t.j);
  t.out_off = net->msg_off[m];
  t.old_off = with_old ? net->msg_off[m] : -1;
  t.normalize = normalize;
  t.slot = slot;
  return t;
}

// Plan a set of open-contraction tasks: order them by path (small LDS classes,
// then large batches) and size the scratch buffers.  Appends to plan.
int plan_tasks(itnbp_net* net, const std::vector<OcTask>& tasks, int acc_f64, TaskPlan& plan) {
  itnbp_ctx* ctx = net->ctx;
  const int eb = acc_f64 ? 8 : 4;
  std::vector<OcTask> small[3], large;
  const size_t lds_cls[3] = {32 * 1024, 64 * 1024, 160 * 1024};
  for (const auto& t : tasks) {
    const size_t b = oc_small_lds_bytes(net->vdesc[t.v], t, eb);
    if (b == 0) {
      large.push_back(t);
      continue;
    }
    const int c = b <= lds_cls[0] ? 0 : b <= lds_cls[1] ? 1 : 2;
    small[c].push_back(t);
  }
  for (int c = 0; c < 3; ++c) {
    if (small[c].empty()) continue;
    LaunchGroup lg{};
    lg.kind = 0;
    lg.begin = plan.tasks.size();
    lg.count = small[c].size();
    for (auto& t : small[c]) lg.lds = std::max(lg.lds, oc_small_lds_bytes(net->vdesc[t.v], t, eb));
    plan.tasks.insert(plan.tasks.end(), small[c].begin(), small[c].end());
    plan.groups.push_back(lg);
  }
  if (!large.empty()) {
    size_t budget = (size_t)6 << 30;
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) == hipSuccess) budget = std::min(budget, freeb / 3);
    std::vector<long long> D(large.size());
    std::vector<int> P(large.size());
    for (size_t i = 0; i < large.size(); ++i) {
      const VDesc& V = net->vdesc[large[i].v];
      D[i] = V.D;
      P[i] = (large[i].s_open ? V.dims[0] : 1) * (large[i].j > 0 ? V.dims[large[i].j] : 1);
    }
    std::vector<itnbp_graph::LargeBatch> batches;
    std::string why;
    if (const int rc = itnbp_graph::large_batches(D, P, budget, eb, batches, why)) return set_err(ctx, rc, why);
    for (const auto& bt : batches) {
      LaunchGroup lg{};
      lg.kind = 1;
      lg.begin = plan.tasks.size();
      lg.count = bt.count;
      lg.maxD = bt.maxD;
      lg.nsplit = std::max(1, std::min(64, bt.maxR / 1024));
      lg.part_stride = (long long)lg.nsplit * bt.maxP2;
      plan.scratch_bytes = std::max(plan.scratch_bytes, lg.count * (size_t)(2 * bt.maxD * eb));
      plan.partial_bytes = std::max(plan.partial_bytes, lg.count * (size_t)lg.part_stride * eb);
      plan.tasks.insert(plan.tasks.end(), large.begin() + bt.begin, large.begin() + bt.begin + bt.count);
      plan.groups.push_back(lg);
    }
  }
  plan.acc_f64 = acc_f64;
  return ITNBP_OK;
}

// upload the task table (synchronously) and size the scratch buffers
int upload_plan(itnbp_net* net, const TaskPlan& plan, DevBuf& taskbuf) {
  itnbp_ctx* ctx = net->ctx;
  if (plan.tasks.empty()) return ITNBP_OK;
  RETCHK(ensure_buf(ctx, taskbuf, plan.tasks.size() * sizeof(OcTask)));
  HIPCHK(ctx, hipMemcpyAsync(taskbuf.p, plan.tasks.data(), plan.tasks.size() * sizeof(OcTask),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (plan.scratch_bytes) RETCHK(ensure_buf(ctx, net->scratch, plan.scratch_bytes));
  if (plan.partial_bytes) RETCHK(ensure_buf(ctx, net->partials, plan.partial_bytes));
  return ITNBP_OK;
}

// launch groups [g0, g1) of an uploaded plan, on the context's stream or on sm
// (then without the per-kernel timing marks, which live on the context's stream)
int launch_plan(itnbp_net* net, const TaskPlan& plan, size_t g0, size_t g1, const OcTask* d_tasks,
                const void* psi, int psi_dtype, const void* msg_in, void* out, MsgStat* d_stats,
                hipStream_t sm) {
  itnbp_ctx* ctx = net->ctx;
  const bool main = sm == nullptr;
  hipStream_t s = main ? ctx->stream : sm;
  for (size_t gi = g0; gi < g1; ++gi) {
    const LaunchGroup& lg = plan.groups[gi];
    const int e0 = main ? prof_mark(ctx) : -1;
    if (lg.kind == 0) {
      HIPCHK(ctx, launch_oc_small(psi_dtype, plan.acc_f64, net->d_vdesc, d_tasks + lg.begin, (int)lg.count,
                                  psi, msg_in, out, d_stats, lg.lds, s));
      if (main) prof_add(ctx, "oc_small", e0, prof_mark(ctx));
    } else {
      HIPCHK(ctx, launch_oc_large(psi_dtype, plan.acc_f64, net->d_vdesc, d_tasks + lg.begin,
                                  plan.tasks.data() + lg.begin, net->vdesc.data(), (int)lg.count, psi,
                                  msg_in, out, d_stats, net->scratch.p, lg.maxD, net->partials.p,
                                  lg.nsplit, lg.part_stride, s));
      if (main) prof_add(ctx, "oc_large", e0, prof_mark(ctx));
    }
  }
  return ITNBP_OK;
}

// plan + upload + launch in one go (environments, scalars, single messages)
int run_tasks(itnbp_net* net, const std::vector<OcTask>& tasks, int acc_f64, const void* psi,
              int psi_dtype, const void* msg_in, void* out, MsgStat* d_stats, DevBuf& taskbuf) {
  if (tasks.empty()) return ITNBP_OK;
  TaskPlan plan;
  RETCHK(plan_tasks(net, tasks, acc_f64, plan));
  RETCHK(upload_plan(net, plan, taskbuf));
  RETCHK(launch_plan(net, plan, 0, plan.groups.size(), (const OcTask*)taskbuf.p, psi, psi_dtype, msg_in,
                     out, d_stats));
  return ITNBP_OK;
}

// ---------------------------------------------------------------------------
// sweep driver
// ---------------------------------------------------------------------------
namespace {

std::vector<std::pair<int, int>> sequential_runs(const itnbp_net* net, const std::vector<int>& seq) {
  return itnbp_graph::sequential_runs(topo_of(net), seq);
}

struct Phase {
  int acc_f64;
  const void* psi;
  int psi_dtype;
  void* msg[2];  // double buffers in the accumulator precision
};

}  // namespace

int sweep_phase(itnbp_net* net, const std::vector<int>& seq, const itnbp_update_opts* o,
                const Phase& ph, int maxiter, double stop_res2, bool stop_on_tol, int* iters,
                double* mean_diff, double* max_res, int* converged, bool use_fast_in,
                double* trace = nullptr, long long trace_stride = 0, int trace_every = 1) {
  bool use_fast = use_fast_in;
  itnbp_ctx* ctx = net->ctx;
  const int n = (int)seq.size();
  RETCHK(ensure_buf(ctx, net->stats, sizeof(MsgStat) * (size_t)std::max(1, n)));
  RETCHK(ensure_buf(ctx, net->red, 8 * sizeof(double)));
  MsgStat* d_stats = (MsgStat*)net->stats.p;
  double* d_red = (double*)net->red.p;
  const size_t eb = ph.acc_f64 ? 8 : 4;
  const bool partial = (o->schedule == ITNBP_SCHED_PARALLEL) && n != 2 * net->ne;
  std::vector<std::pair<int, int>> runs;
  if (o->schedule == ITNBP_SCHED_PARALLEL) {
    runs.push_back({0, n});
  } else {
    runs = sequential_runs(net, seq);
  }
  // the specialised pipeline produces the messages of its bulk vertices; the
  // generic plan everything else
  std::vector<char> handled(n, 0);
  if (use_fast) {
    const int fr = fast_prepare(net, seq, ph.acc_f64, handled);
    if (fr == ITNBP_ERR_UNSUPPORTED) {
      use_fast = false;
      std::fill(handled.begin(), handled.end(), 0);
    } else if (fr != ITNBP_OK) {
      return fr;
    }
  }
  // whole-vertex tasks for the small-chi vertices the pipeline did not take
  bool use_small = false;
  if (o->schedule == ITNBP_SCHED_PARALLEL) {
    RETCHK(small_prepare(net, seq, ph.acc_f64, handled));
    use_small = small_task_count(net) > 0;
  }
  // one plan for the whole sweep; run r launches groups [run_groups[r], run_groups[r+1])
  TaskPlan plan;
  std::vector<size_t> run_groups{0};
  for (auto& r : runs) {
    std::vector<OcTask> rt;
    rt.reserve(r.second - r.first);
    for (int i = r.first; i < r.second; ++i)
      if (!handled[i]) rt.push_back(message_task(net, seq[i], i, true, o->normalize));
    RETCHK(plan_tasks(net, rt, ph.acc_f64, plan));
    run_groups.push_back(plan.groups.size());
  }
  RETCHK(upload_plan(net, plan, net->tasks));
  const OcTask* d_tasks = (const OcTask*)net->tasks.p;
  int it = 0;
  *converged = 0;
  double red[4] = {0, 0, 0, 0};
  double best_res2 = INFINITY;
  int stall = 0;
  const bool watch = stop_res2 >= 0 || (stop_on_tol && (o->tol >= 0 || o->res_tol >= 0)) || o->verbose;
  for (it = 1; it <= maxiter; ++it) {
    if (o->schedule == ITNBP_SCHED_PARALLEL) {
      void* cur = ph.msg[net->cur];
      void* nxt = ph.msg[1 - net->cur];
      if (partial)
        HIPCHK(ctx, hipMemcpyAsync(nxt, cur, net->msg_total * eb, hipMemcpyDeviceToDevice, ctx->stream));
      // beside the chi = 32 pipeline, the generic tasks (e.g. the lattice corners)
      // run on the side stream: they read cur and write disjoint messages and
      // statistics slots of their own
      const bool fork = use_fast && !plan.groups.empty();
      if (fork) {
        HIPCHK(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
        HIPCHK(ctx, hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        RETCHK(launch_plan(net, plan, 0, plan.groups.size(), d_tasks, ph.psi, ph.psi_dtype, cur, nxt, d_stats,
                           ctx->side));
        HIPCHK(ctx, hipEventRecord(ctx->ev_join, ctx->side));
      }
      if (use_fast) RETCHK(fast_sweep(net, ph.acc_f64, cur, nxt, d_stats, o->normalize));
      if (use_small) RETCHK(small_sweep(net, ph.acc_f64, ph.psi, ph.psi_dtype, cur, nxt, d_stats, o->normalize));
      if (fork)
        HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
      else
        RETCHK(launch_plan(net, plan, 0, plan.groups.size(), d_tasks, ph.psi, ph.psi_dtype, cur, nxt, d_stats));
      net->cur = 1 - net->cur;
    } else {
      void* cur = ph.msg[net->cur];
      for (size_t r = 0; r + 1 < run_groups.size(); ++r)
        RETCHK(launch_plan(net, plan, run_groups[r], run_groups[r + 1], d_tasks, ph.psi, ph.psi_dtype, cur, cur, d_stats));
    }
    // FP64 messages after every trace_every-th sweep to trace + (it / trace_every) trace_stride
    if (trace && it % trace_every == 0)
      HIPCHK(ctx, hipMemcpyAsync(trace + (long long)(it / trace_every) * trace_stride, ph.msg[net->cur],
                                 net->msg_total * 8, hipMemcpyDeviceToDevice, ctx->stream));
    // with no stopping or hand-over rule the statistics are read after the last
    // sweep only (a non-finite message persists through later sweeps, so it is
    // still caught there)
    if (!watch && it < maxiter) continue;
    HIPCHK(ctx, launch_reduce_stats(d_stats, n, d_red, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(red, d_red, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (red[2] >= 1.0)
      return set_err(ctx, ITNBP_ERR_NONFINITE, "non-finite message after sweep " + std::to_string(it));
    *mean_diff = red[0] / n;
    *max_res = std::sqrt(red[1]);
    bool stop = false;
    if (stop_on_tol && o->tol >= 0 && *mean_diff <= o->tol) stop = true;
    if (stop_on_tol && o->res_tol >= 0 && red[1] <= o->res_tol * o->res_tol) stop = true;
    if (stop_res2 >= 0 && red[1] <= stop_res2) stop = true;
    // reduced-precision phase: also hand over once the residual stagnates at the
    // precision floor (no 20% improvement over two consecutive sweeps)
    if (stop_res2 >= 0) {
      if (red[1] < 0.64 * best_res2) {
        best_res2 = red[1];
        stall = 0;
      } else if (++stall >= 2 && it >= 3) {
        stop = true;
      }
    }
    if (stop) {
      *converged = 1;
      if (o->verbose) printf("BP converged to desired precision after %d iterations.\n", it);
      break;
    }
  }
  *iters = std::min(it, maxiter);
  return ITNBP_OK;
}

namespace {

int default_sequence(const itnbp_net* net, int schedule, std::vector<int>& seq);

// reduced-precision modes (every ITNBP_PREC_* but F64)
bool prec_known(int p) { return p >= ITNBP_PREC_F32 && p <= ITNBP_PREC_BF16_F16X2; }
// compensated arithmetic of the (last) reduced-precision phase: 0 none, 1 three
// BF16 terms, 2 two FP16 terms
int prec_split(int p) {
  if (p == ITNBP_PREC_F16X2 || p == ITNBP_PREC_BF16_F16X2) return 2;
  if (p == ITNBP_PREC_BF16X3 || p == ITNBP_PREC_BF16X3_F64 || p == ITNBP_PREC_BF16_BF16X3 ||
      p == ITNBP_PREC_BF16_BF16X3_F64)
    return 1;
  return 0;
}
// one-term 16-bit phase first, then the compensated one
bool prec_progressive(int p) {
  return p == ITNBP_PREC_BF16_BF16X3 || p == ITNBP_PREC_BF16_BF16X3_F64 || p == ITNBP_PREC_BF16_F16X2;
}
// no FP64 refinement phase
bool prec_low_only(int p) {
  return p == ITNBP_PREC_F32 || p == ITNBP_PREC_BF16X3 || p == ITNBP_PREC_BF16_BF16X3 || p == ITNBP_PREC_F16X2 ||
         p == ITNBP_PREC_BF16_F16X2;
}

int check_sequence(itnbp_net* net, const std::vector<int>& seq, const itnbp_update_opts* o) {
  // every phased mode measures the residual against the previous messages
  const bool need_old = o->tol >= 0 || o->res_tol >= 0 ||
                        (o->precision != ITNBP_PREC_F64 && o->precision != ITNBP_PREC_F32 &&
                         o->precision != ITNBP_PREC_BF16X3 && o->precision != ITNBP_PREC_F16X2);
  std::string why;
  const int rc = itnbp_graph::check_sequence(topo_of(net), net->msg_set, seq, o->schedule == ITNBP_SCHED_SEQUENTIAL,
                                             need_old, why);
  return rc ? set_err(net->ctx, rc, why) : ITNBP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// nsweeps FP64 Jacobi sweeps over the default parallel schedule, messages
// normalised, no convergence test (forward pass of the adjoint); with trace,
// the messages after every trace_every-th sweep it go to
// trace + (it / trace_every) trace_stride
int sweep_phase_public(itnbp_net* net, int nsweeps, double* trace, long long trace_stride, int trace_every) {
  std::vector<int> seq;
  RETCHK(default_sequence(net, ITNBP_SCHED_PARALLEL, seq));
  if (seq.empty()) return ITNBP_OK;
  itnbp_update_opts o{};
  o.schedule = ITNBP_SCHED_PARALLEL;
  o.maxiter = nsweeps;
  o.tol = -1;
  o.res_tol = -1;
  o.precision = ITNBP_PREC_F64;
  o.normalize = 1;
  Phase ph{1, net->d_psi, net->dtype, {net->d_msg[0], net->d_msg[1]}};
  int it = 0, conv = 0;
  double md = 0, mr = 0;
  return sweep_phase(net, seq, &o, ph, nsweeps, -1.0, false, &it, &md, &mr, &conv,
                     fast_path_available(net, 1, ITNBP_SCHED_PARALLEL), trace, trace_stride, trace_every);
}

extern "C" {

int itnbp_create(int device, itnbp_ctx** out) {
  if (!out) return ITNBP_ERR_ARG;
  *out = nullptr;
  auto* ctx = new (std::nothrow) itnbp_ctx();
  if (!ctx) return ITNBP_ERR_NOMEM;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || device < 0 || device >= n) {
    (void)hipGetLastError();
    delete ctx;
    return ITNBP_ERR_HIP;
  }
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipStreamCreate(&ctx->side) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
    delete ctx;
    return ITNBP_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
  // the workspace of itnbp_apply_gates_noalloc on this device (a failure here
  // only disables that entry point, which then reports it)
  (void)apply_noalloc_init(device);
  *out = ctx;
  return ITNBP_OK;
}

int itnbp_destroy(itnbp_ctx* ctx) {
  if (!ctx) return ITNBP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return ITNBP_OK;
}

const char* itnbp_last_error(itnbp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int itnbp_profile_enable(itnbp_ctx* ctx, int enable) {
  if (!ctx) return ITNBP_ERR_ARG;
  ctx->prof = enable != 0;
  return ITNBP_OK;
}

int itnbp_profile_reset(itnbp_ctx* ctx) {
  if (!ctx) return ITNBP_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->recs.clear();
  ctx->ev_used = 0;
  return ITNBP_OK;
}

int itnbp_profile_query(itnbp_ctx* ctx, const char* name, int* count, double* total_ms) {
  if (!ctx || !name || !count || !total_ms) return ITNBP_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  int c = 0;
  double t = 0;
  for (const auto& r : ctx->recs) {
    if (strcmp(r.name, name) != 0) continue;
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev_pool[r.ev0], ctx->ev_pool[r.ev1]));
    ++c;
    t += ms;
  }
  *count = c;
  *total_ms = t;
  return ITNBP_OK;
}
const char* itnbp_last_path(itnbp_ctx* ctx) { return ctx ? ctx->path.c_str() : ""; }

int itnbp_synchronize(itnbp_ctx* ctx) {
  if (!ctx) return ITNBP_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

int itnbp_net_create(itnbp_ctx* ctx, int nv, int ne, const int* edges, const int* inc_ptr,
                     const int* inc_edge, const int* bond_dims, const int* phys_dims, int dtype,
                     itnbp_net** out) {
  if (!ctx || !out) return ITNBP_ERR_ARG;
  *out = nullptr;
  return guarded(ctx, [&]() -> int {
    if (dtype != ITNBP_F64 && dtype != ITNBP_F32 && dtype != ITNBP_BF16)
      return set_err(ctx, ITNBP_ERR_ARG, "unknown dtype");
    const itnbp_graph::Topo topo{nv, ne, edges, inc_ptr, inc_edge};
    const itnbp_graph::Limits lim{ITNBP_MAXK, 64, 8, 1LL << 30};
    itnbp_graph::Layout lay;
    std::string why;
    if (const int rc = itnbp_graph::describe(topo, bond_dims, phys_dims, lim, lay, why)) return set_err(ctx, rc, why);
    auto* net = new itnbp_net();
    std::unique_ptr<itnbp_net> guard(net);
    net->ctx = ctx;
    net->nv = nv;
    net->ne = ne;
    net->dtype = dtype;
    net->edges.assign(edges, edges + 2 * ne);
    net->inc_ptr.assign(inc_ptr, inc_ptr + nv + 1);
    net->inc_edge.assign(inc_edge, inc_edge + 2 * ne);
    net->chi.assign(bond_dims, bond_dims + ne);
    net->phys.assign(phys_dims, phys_dims + nv);
    net->msg_off = lay.msg_off;
    net->msg_total = lay.msg_total;
    net->max_chi = lay.max_chi;
    net->max_k = lay.max_k;
    net->msg_set.assign(2 * ne, 0);
    net->psi_off = lay.psi_off;
    net->psi_total = lay.psi_total;
    net->vdesc.resize(nv);
    for (int v = 0; v < nv; ++v) {
      VDesc& V = net->vdesc[v];
      memset(&V, 0, sizeof(V));
      V.k = net->inc_ptr[v + 1] - net->inc_ptr[v];
      V.dims[0] = net->phys[v];
      for (int a = 1; a <= V.k; ++a) {
        const int e = net->inc_edge[net->inc_ptr[v] + a - 1];
        V.dims[a] = net->chi[e];
        V.in_msg[a] = net->msg_off[msg_id_into(net, e, v)];
      }
      V.D = (int)lay.site_D[v];
      V.psi_off = lay.psi_off[v];
    }
    const long long po = lay.psi_total, mo = lay.msg_total;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    if (hipMalloc(&net->d_psi, std::max<size_t>(16, po * dtype_size(dtype))) != hipSuccess ||
        hipMalloc((void**)&net->d_msg[0], std::max<size_t>(16, mo * 8)) != hipSuccess ||
        hipMalloc((void**)&net->d_msg[1], std::max<size_t>(16, mo * 8)) != hipSuccess ||
        hipMalloc((void**)&net->d_vdesc, sizeof(VDesc) * nv) != hipSuccess) {
      (void)hipGetLastError();
      itnbp_net_destroy(guard.release());
      return set_err(ctx, ITNBP_ERR_NOMEM, "device allocation for the network failed");
    }
    HIPCHK(ctx, hipMemcpyAsync(net->d_vdesc, net->vdesc.data(), sizeof(VDesc) * nv, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(net->d_psi, 0, po * dtype_size(dtype), ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(net->d_msg[0], 0, mo * 8, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *out = guard.release();
    return ITNBP_OK;
  });
}

int itnbp_net_destroy(itnbp_net* net) {
  if (!net) return ITNBP_OK;
  if (net->ctx) {
    (void)hipSetDevice(net->ctx->device);
    (void)hipStreamSynchronize(net->ctx->stream);
  }
  if (net->d_psi) (void)hipFree(net->d_psi);
  if (net->d_psi64) (void)hipFree(net->d_psi64);
  for (int i = 0; i < 2; ++i) {
    if (net->d_msg[i]) (void)hipFree(net->d_msg[i]);
    if (net->d_msg32[i]) (void)hipFree(net->d_msg32[i]);
  }
  if (net->d_vdesc) (void)hipFree(net->d_vdesc);
  free_buf(net->scratch);
  free_buf(net->tasks);
  free_buf(net->stats);
  free_buf(net->partials);
  free_buf(net->red);
  free_buf(net->aux);
  free_buf(net->mdesc);
  free_buf(net->apply_ws);
  release_fast_state(net);
  release_small_state(net);
  delete net;
  return ITNBP_OK;
}

int itnbp_net_copy(itnbp_net* src, itnbp_net** out) {
  if (!src || !out) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = src->ctx;
  RETCHK(itnbp_net_create(ctx, src->nv, src->ne, src->edges.data(), src->inc_ptr.data(),
                          src->inc_edge.data(), src->chi.data(), src->phys.data(), src->dtype, out));
  itnbp_net* dst = *out;
  HIPCHK(ctx, hipMemcpyAsync(dst->d_psi, src->d_psi, src->psi_total * dtype_size(src->dtype),
                             hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(dst->d_msg[0], src->d_msg[src->cur], src->msg_total * 8,
                             hipMemcpyDeviceToDevice, ctx->stream));
  dst->cur = 0;
  dst->msg_set = src->msg_set;
  dst->msgs_sym = src->msgs_sym;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

}  // extern "C"

// ---- bond resizing -------------------------------------------------------------
namespace {

struct RepackDesc {
  long long src_off, dst_off, D;      // element offsets and new element count
  int r;                              // rank (physical + bonds)
  int old_dims[ITNBP_MAXR], new_dims[ITNBP_MAXR];
};

// dst tensor v = src tensor v cut or zero-padded along every axis to the new
// dimensions (column-major); grid (chunks, vertices)
template <typename T>
__global__ void k_repack_psi(const T* __restrict__ src, T* __restrict__ dst, const RepackDesc* __restrict__ rd) {
  const RepackDesc d = rd[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < d.D; i += (long long)gridDim.x * blockDim.x) {
    long long rem = i, so = 0, stride = 1;
    bool in = true;
    for (int a = 0; a < d.r; ++a) {
      const int x = (int)(rem % d.new_dims[a]);
      rem /= d.new_dims[a];
      in = in && x < d.old_dims[a];
      so += stride * x;
      stride *= d.old_dims[a];
    }
    T z;
    memset(&z, 0, sizeof(T));
    dst[d.dst_off + i] = in ? src[d.src_off + so] : z;
  }
}

struct MsgRepack {
  long long src_off, dst_off;
  int chi, reset;  // new chi; reset: identity of the new dimension
};

__global__ void k_repack_msgs(const double* __restrict__ src, double* __restrict__ dst,
                              const MsgRepack* __restrict__ mr) {
  const MsgRepack m = mr[blockIdx.x];
  for (int i = threadIdx.x; i < m.chi * m.chi; i += blockDim.x)
    dst[m.dst_off + i] = m.reset ? ((i % m.chi) == (i / m.chi) ? 1.0 : 0.0) : src[m.src_off + i];
}

}  // namespace

// Gives every edge the bond dimension chi_new[e]: site tensors are cut or
// zero-padded along the changed bonds, the messages of a changed edge become
// identities of the new dimension (set), the others are kept.  Every state
// derived from the bond dimensions is released.
int net_resize_bonds(itnbp_net* net, const std::vector<int>& chi_new) {
  itnbp_ctx* ctx = net->ctx;
  if ((int)chi_new.size() != net->ne) return set_err(ctx, ITNBP_ERR_ARG, "resize: one dimension per edge");
  if (chi_new == net->chi) return ITNBP_OK;
  itnbp_net* tmp = nullptr;
  RETCHK(itnbp_net_create(ctx, net->nv, net->ne, net->edges.data(), net->inc_ptr.data(), net->inc_edge.data(),
                          chi_new.data(), net->phys.data(), net->dtype, &tmp));
  std::unique_ptr<itnbp_net, int (*)(itnbp_net*)> guard(tmp, itnbp_net_destroy);
  std::vector<RepackDesc> rd(net->nv);
  for (int v = 0; v < net->nv; ++v) {
    const VDesc& A = net->vdesc[v];
    const VDesc& B = tmp->vdesc[v];
    rd[v].src_off = A.psi_off;
    rd[v].dst_off = B.psi_off;
    rd[v].D = B.D;
    rd[v].r = A.k + 1;
    for (int a = 0; a <= A.k; ++a) {
      rd[v].old_dims[a] = A.dims[a];
      rd[v].new_dims[a] = B.dims[a];
    }
  }
  std::vector<MsgRepack> mr(2 * net->ne);
  for (int m = 0; m < 2 * net->ne; ++m) {
    const int e = m / 2;
    mr[m].src_off = net->msg_off[m];
    mr[m].dst_off = tmp->msg_off[m];
    mr[m].chi = chi_new[e];
    mr[m].reset = chi_new[e] != net->chi[e] ? 1 : 0;
  }
  DevBuf b1, b2;
  RETCHK(upload(ctx, b1, rd.data(), rd.size() * sizeof(RepackDesc)));
  if (!mr.empty()) RETCHK(upload(ctx, b2, mr.data(), mr.size() * sizeof(MsgRepack)));
  long long maxD = 1;
  for (auto& d : rd) maxD = std::max(maxD, d.D);
  const dim3 grid((unsigned)std::min<long long>((maxD + 255) / 256, 1024), net->nv);
  switch (net->dtype) {
    case ITNBP_F64:
      hipLaunchKernelGGL(k_repack_psi<double>, grid, dim3(256), 0, ctx->stream, (const double*)net->d_psi,
                         (double*)tmp->d_psi, (const RepackDesc*)b1.p);
      break;
    case ITNBP_F32:
      hipLaunchKernelGGL(k_repack_psi<float>, grid, dim3(256), 0, ctx->stream, (const float*)net->d_psi,
                         (float*)tmp->d_psi, (const RepackDesc*)b1.p);
      break;
    default:
      hipLaunchKernelGGL(k_repack_psi<bf16_t>, grid, dim3(256), 0, ctx->stream, (const bf16_t*)net->d_psi,
                         (bf16_t*)tmp->d_psi, (const RepackDesc*)b1.p);
      break;
  }
  if (!mr.empty())
    hipLaunchKernelGGL(k_repack_msgs, dim3(2 * net->ne), dim3(256), 0, ctx->stream, (const double*)net->d_msg[net->cur],
                       tmp->d_msg[0], (const MsgRepack*)b2.p);
  hipError_t ek = hipGetLastError();
  if (ek == hipSuccess) ek = hipStreamSynchronize(ctx->stream);
  free_buf(b1);
  free_buf(b2);
  if (ek != hipSuccess) return set_err(ctx, ITNBP_ERR_HIP, std::string("resize: ") + hipGetErrorString(ek));
  // hand tmp's description and buffers to net; net's old ones go with tmp
  release_fast_state(net);
  release_small_state(net);
  for (int m = 0; m < 2 * net->ne; ++m)
    if (mr[m].reset) net->msg_set[m] = 1;
  std::swap(net->chi, tmp->chi);
  std::swap(net->psi_off, tmp->psi_off);
  std::swap(net->msg_off, tmp->msg_off);
  std::swap(net->psi_total, tmp->psi_total);
  std::swap(net->msg_total, tmp->msg_total);
  std::swap(net->vdesc, tmp->vdesc);
  std::swap(net->max_chi, tmp->max_chi);
  std::swap(net->max_k, tmp->max_k);
  std::swap(net->d_psi, tmp->d_psi);
  std::swap(net->d_psi64, tmp->d_psi64);
  for (int i = 0; i < 2; ++i) {
    std::swap(net->d_msg[i], tmp->d_msg[i]);
    std::swap(net->d_msg32[i], tmp->d_msg32[i]);
  }
  std::swap(net->d_vdesc, tmp->d_vdesc);
  std::swap(net->scratch, tmp->scratch);
  std::swap(net->tasks, tmp->tasks);
  std::swap(net->stats, tmp->stats);
  std::swap(net->partials, tmp->partials);
  std::swap(net->red, tmp->red);
  std::swap(net->aux, tmp->aux);
  std::swap(net->mdesc, tmp->mdesc);
  net->cur = 0;
  return ITNBP_OK;
}

extern "C" {

int itnbp_get_bond_dims(itnbp_net* net, int* bond_dims) {
  if (!net || (net->ne > 0 && !bond_dims)) return ITNBP_ERR_ARG;
  std::copy(net->chi.begin(), net->chi.end(), bond_dims);
  return ITNBP_OK;
}

int64_t itnbp_net_psi_numel(itnbp_net* net) { return net ? net->psi_total : -1; }
int64_t itnbp_net_msg_numel(itnbp_net* net) { return net ? net->msg_total : -1; }

int itnbp_set_sites(itnbp_net* net, const void* data) {
  if (!net || !data) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  HIPCHK(ctx, hipMemcpyAsync(net->d_psi, data, net->psi_total * dtype_size(net->dtype),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  invalidate_fast_state(net);
  return ITNBP_OK;
}

int itnbp_get_sites(itnbp_net* net, void* data) {
  if (!net || !data) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  HIPCHK(ctx, hipMemcpyAsync(data, net->d_psi, net->psi_total * dtype_size(net->dtype),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

int itnbp_set_site(itnbp_net* net, int v, const void* data) {
  if (!net || !data || v < 0 || v >= net->nv) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  const size_t eb = dtype_size(net->dtype);
  HIPCHK(ctx, hipMemcpyAsync((char*)net->d_psi + net->psi_off[v] * eb, data, net->vdesc[v].D * eb,
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  invalidate_fast_state(net);
  return ITNBP_OK;
}

// symmetric to within rounding: max |M - M^T| <= 1e-12 max |M|
static bool message_symmetric(const double* M, int chi) {
  double mx = 0, d = 0;
  for (int i = 0; i < chi; ++i)
    for (int j = 0; j < chi; ++j) {
      mx = std::max(mx, std::fabs(M[i + chi * j]));
      d = std::max(d, std::fabs(M[i + chi * j] - M[j + chi * i]));
    }
  return d <= 1e-12 * mx;
}

int itnbp_set_messages(itnbp_net* net, const double* data) {
  if (!net || !data) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  HIPCHK(ctx, hipMemcpyAsync(net->d_msg[net->cur], data, net->msg_total * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::fill(net->msg_set.begin(), net->msg_set.end(), 1);
  bool sym = true;
  for (int m = 0; m < 2 * net->ne && sym; ++m) sym = message_symmetric(data + net->msg_off[m], net->chi[m / 2]);
  net->msgs_sym = sym;
  return ITNBP_OK;
}

int itnbp_get_messages(itnbp_net* net, double* data) {
  if (!net || !data) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  HIPCHK(ctx, hipMemcpyAsync(data, net->d_msg[net->cur], net->msg_total * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

int itnbp_set_message(itnbp_net* net, int m, const double* data) {
  if (!net || !data || m < 0 || m >= 2 * net->ne) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  const long long n = (long long)net->chi[m / 2] * net->chi[m / 2];
  HIPCHK(ctx, hipMemcpyAsync(net->d_msg[net->cur] + net->msg_off[m], data, n * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  net->msg_set[m] = 1;
  if (!message_symmetric(data, net->chi[m / 2])) net->msgs_sym = false;
  return ITNBP_OK;
}

int itnbp_get_message(itnbp_net* net, int m, double* data) {
  if (!net || !data || m < 0 || m >= 2 * net->ne) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  if (!net->msg_set[m]) return set_err(ctx, ITNBP_ERR_STATE, "message " + std::to_string(m) + " is not set");
  const long long n = (long long)net->chi[m / 2] * net->chi[m / 2];
  HIPCHK(ctx, hipMemcpyAsync(data, net->d_msg[net->cur] + net->msg_off[m], n * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return ITNBP_OK;
}

int itnbp_message_is_set(itnbp_net* net, int m) {
  if (!net || m < 0 || m >= 2 * net->ne) return 0;
  return net->msg_set[m] ? 1 : 0;
}

namespace {
// every message M_m (chi_m x chi_m at offset off_m) <- identity; one workgroup per message
__global__ void k_identity_msgs(double* __restrict__ msg, const long long* __restrict__ off,
                                const int* __restrict__ chi, int nmsg) {
  const int m = blockIdx.x;
  if (m >= nmsg) return;
  const int c = chi[m];
  double* M = msg + off[m];
  for (int t = threadIdx.x; t < c * c; t += blockDim.x) M[t] = (t % (c + 1) == 0) ? 1.0 : 0.0;
}
}  // namespace

// identity_messages (quadraticformnetwork.jl:96-124), written on the device
int itnbp_init_identity_messages(itnbp_net* net) {
  if (!net) return ITNBP_ERR_ARG;
  return guarded(net->ctx, [&]() -> int {
    itnbp_ctx* ctx = net->ctx;
    const int nmsg = 2 * net->ne;
    if (nmsg > 0) {
      const size_t bytes = (size_t)nmsg * (sizeof(long long) + sizeof(int));
      if (net->mdesc.bytes < bytes) {
        std::vector<char> h(bytes);
        long long* off = reinterpret_cast<long long*>(h.data());
        int* chi = reinterpret_cast<int*>(off + nmsg);
        for (int m = 0; m < nmsg; ++m) {
          off[m] = net->msg_off[m];
          chi[m] = net->chi[m / 2];
        }
        RETCHK(ensure_buf(ctx, net->mdesc, bytes));
        HIPCHK(ctx, hipMemcpyAsync(net->mdesc.p, h.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      }
      const long long* d_off = (const long long*)net->mdesc.p;
      const int* d_chi = (const int*)(d_off + nmsg);
      hipLaunchKernelGGL(k_identity_msgs, dim3(nmsg), dim3(256), 0, ctx->stream, net->d_msg[net->cur], d_off, d_chi,
                         nmsg);
      HIPCHK(ctx, hipGetLastError());
    }
    std::fill(net->msg_set.begin(), net->msg_set.end(), 1);
    net->msgs_sym = true;
    return ITNBP_OK;
  });
}

int itnbp_forest_cover_sequence(itnbp_net* net, int* out) {
  if (!net || !out) return ITNBP_ERR_ARG;
  return guarded(net->ctx, [&]() -> int {
    std::vector<int> seq;
    RETCHK(default_sequence(net, ITNBP_SCHED_SEQUENTIAL, seq));
    std::copy(seq.begin(), seq.end(), out);
    return ITNBP_OK;
  });
}

int itnbp_update(itnbp_net* net, const itnbp_update_opts* o, itnbp_update_result* res) {
  if (!net || !o) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  return guarded(ctx, [&]() -> int {
    if (o->maxiter < 1) return set_err(ctx, ITNBP_ERR_ARG, "You need to specify a number of iterations for BP!");
    if (o->schedule != ITNBP_SCHED_SEQUENTIAL && o->schedule != ITNBP_SCHED_PARALLEL)
      return set_err(ctx, ITNBP_ERR_ARG, "unknown schedule");
    std::vector<int> seq;
    if (o->nseq > 0) {
      if (!o->seq) return set_err(ctx, ITNBP_ERR_ARG, "nseq > 0 with a null seq");
      seq.assign(o->seq, o->seq + o->nseq);
    } else {
      RETCHK(default_sequence(net, o->schedule, seq));
    }
    if (seq.empty()) {
      if (res) *res = itnbp_update_result{0, 0, 0.0, 0.0, 1, 0.0};
      return ITNBP_OK;
    }
    RETCHK(check_sequence(net, seq, o));
    HIPCHK(ctx, hipSetDevice(ctx->device));
    itnbp_update_result r{};
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    int it = 0, conv = 0;
    double md = 0, mr = 0;
    if (o->precision == ITNBP_PREC_F64) {
      Phase ph{1, net->d_psi, net->dtype, {net->d_msg[0], net->d_msg[1]}};
      const bool fast = fast_path_available(net, 1, o->schedule);
      ctx->path = fast ? fast_path_name(net, 1) : "generic-f64";
      RETCHK(sweep_phase(net, seq, o, ph, o->maxiter, -1.0, true, &it, &md, &mr, &conv, fast));
      r.iterations = it;
    } else if (prec_known(o->precision)) {
      const int split = prec_split(o->precision);
      const bool progressive = prec_progressive(o->precision);
      for (int i = 0; i < 2; ++i)
        if (!net->d_msg32[i]) HIPCHK(ctx, hipMalloc((void**)&net->d_msg32[i], std::max<size_t>(16, net->msg_total * 4)));
      HIPCHK(ctx, launch_convert_msgs(net->d_msg[net->cur], net->d_msg32[net->cur], net->msg_total, ctx->stream));
      Phase ph{0, net->d_psi, net->dtype, {net->d_msg32[0], net->d_msg32[1]}};
      const bool low_only = prec_low_only(o->precision);
      const double sw2 = low_only ? -1.0 : o->switch_tol * o->switch_tol;
      // progressive: the one-term 16-bit kernels until the residual < switch_tol
      // (or stagnates), then the three-term kernels
      int it_pre = 0;
      ctx->path.clear();
      if (progressive) {
        net->low_x3 = 0;
        const bool fast0 = fast_path_available(net, 0, o->schedule);
        ctx->path = std::string(fast0 ? fast_path_name(net, 0) : "generic-f32") + "+";
        // BF16_BF16X3: switch_tol ends the one-term phase; BF16_BF16X3_F64: the
        // one-term phase ends at residual 1e-2 (two decades above its floor) and
        // switch_tol ends the three-term phase
        const double pre = o->precision == ITNBP_PREC_BF16_BF16X3_F64 ? 1e-2 : o->switch_tol;
        const double pre2 = pre > 0 ? pre * pre : 0.0;
        const int rc_pre = sweep_phase(net, seq, o, ph, o->maxiter, pre2, false, &it_pre, &md, &mr, &conv, fast0);
        RETCHK(rc_pre);
        conv = 0;
      }
      net->low_x3 = split;
      const bool fast = fast_path_available(net, 0, o->schedule);
      ctx->path += fast ? fast_path_name(net, 0) : "generic-f32";
      int rc_low = ITNBP_OK;
      if (o->maxiter - it_pre > 0)
        rc_low = sweep_phase(net, seq, o, ph, o->maxiter - it_pre, sw2, low_only, &it, &md, &mr, &conv, fast);
      it += it_pre;
      net->low_x3 = 0;
      RETCHK(rc_low);
      HIPCHK(ctx, launch_convert_msgs_back(net->d_msg32[net->cur], net->d_msg[net->cur], net->msg_total, ctx->stream));
      r.iterations_low = it;
      r.iterations = it;
      if (!low_only) {
        conv = 0;
        const int left = o->maxiter - it;
        if (left > 0) {
          Phase ph2{1, net->d_psi, net->dtype, {net->d_msg[0], net->d_msg[1]}};
          const bool fast2 = fast_path_available(net, 1, o->schedule);
          ctx->path += std::string("+") + (fast2 ? fast_path_name(net, 1) : "generic-f64");
          int it2 = 0;
          RETCHK(sweep_phase(net, seq, o, ph2, left, -1.0, true, &it2, &md, &mr, &conv, fast2));
          r.iterations += it2;
        }
      }
    } else {
      return set_err(ctx, ITNBP_ERR_ARG, "unknown precision mode");
    }
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    r.device_ms = ms;
    if (small_task_count(net) > 0) ctx->path += " [vertex " + std::to_string(small_task_count(net)) + "]";
    for (int a = 0; a < 2; ++a)
      if (fast_chunk(net, a) > 0) ctx->path += std::string(a ? " [f64 chunk " : " [bf16 chunk ") + std::to_string(fast_chunk(net, a)) + "]";
    r.mean_diff = md;
    r.max_residual = mr;
    r.converged = conv;
    for (int m : seq) net->msg_set[m] = 1;
    if (res) *res = r;
    return ITNBP_OK;
  });
}

int itnbp_updated_messages(itnbp_net* net, int n, const int* ids, int normalize, double* out) {
  if (!net || n < 0 || (n > 0 && (!ids || !out))) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  return guarded(ctx, [&]() -> int {
    if (n == 0) return ITNBP_OK;
    std::vector<OcTask> tasks;
    long long tot = 0;
    std::vector<long long> offs;
    for (int i = 0; i < n; ++i) {
      if (ids[i] < 0 || ids[i] >= 2 * net->ne) return set_err(ctx, ITNBP_ERR_ARG, "message id out of range");
      OcTask t = message_task(net, ids[i], i, false, normalize);
      t.out_off = tot;
      offs.push_back(tot);
      tot += (long long)net->chi[ids[i] / 2] * net->chi[ids[i] / 2];
      tasks.push_back(t);
    }
    itnbp_update_opts o{};
    o.tol = -1;
    o.res_tol = -1;
    o.schedule = ITNBP_SCHED_PARALLEL;
    RETCHK(check_sequence(net, std::vector<int>(ids, ids + n), &o));
    RETCHK(ensure_buf(ctx, net->aux, tot * 8));
    RETCHK(ensure_buf(ctx, net->stats, sizeof(MsgStat) * n));
    RETCHK(run_tasks(net, tasks, 1, net->d_psi, net->dtype, net->d_msg[net->cur], net->aux.p,
                     (MsgStat*)net->stats.p, net->tasks));
    HIPCHK(ctx, hipMemcpyAsync(out, net->aux.p, tot * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return ITNBP_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// default schedules
// ---------------------------------------------------------------------------
namespace {

int default_sequence(const itnbp_net* net, int schedule, std::vector<int>& seq) {
  itnbp_graph::default_sequence(topo_of(net), schedule == ITNBP_SCHED_PARALLEL, seq);
  return ITNBP_OK;
}

}  // namespace
