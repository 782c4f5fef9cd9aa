// Host-side helpers shared by the libitnbp translation units.
#pragma once
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "itnbp_internal.hpp"

int itnbp_fail(itnbp_ctx* ctx, int code, const std::string& msg);
int ensure_buf(itnbp_ctx* ctx, DevBuf& b, size_t bytes);
void free_buf(DevBuf& b);
int upload(itnbp_ctx* ctx, DevBuf& b, const void* src, size_t bytes);

int msg_id_into(const itnbp_net* net, int e, int to);
int msg_src(const itnbp_net* net, int m);
int msg_dst(const itnbp_net* net, int m);
int axis_of(const itnbp_net* net, int v, int e);
int all_bonds_mask(const itnbp_net* net, int v);
OcTask message_task(const itnbp_net* net, int m, int slot, bool with_old, int normalize);

int plan_tasks(itnbp_net* net, const std::vector<OcTask>& tasks, int acc_f64, TaskPlan& plan);
int upload_plan(itnbp_net* net, const TaskPlan& plan, DevBuf& taskbuf);
int launch_plan(itnbp_net* net, const TaskPlan& plan, size_t g0, size_t g1, const OcTask* d_tasks,
                const void* psi, int psi_dtype, const void* msg_in, void* out, MsgStat* d_stats,
                hipStream_t sm = nullptr);
// run open-contraction tasks; msg_in/out in the accumulator precision
int run_tasks(itnbp_net* net, const std::vector<OcTask>& tasks, int acc_f64, const void* psi,
              int psi_dtype, const void* msg_in, void* out, MsgStat* d_stats, DevBuf& taskbuf);

// ---- specialised sweep kernels (k_fast*.hip) -------------------------------
// true when the specialised parallel-sweep pipeline handles this network
bool fast_path_available(itnbp_net* net, int acc_f64, int schedule);
const char* fast_path_name(itnbp_net* net, int acc_f64);
// mark the entries of seq the pipeline produces and build its buffers
int fast_prepare(itnbp_net* net, const std::vector<int>& seq, int acc_f64, std::vector<char>& handled);
// launch the pipeline for one Jacobi sweep from msg_in into msg_out
int fast_sweep(itnbp_net* net, int acc_f64, const void* msg_in, void* msg_out, MsgStat* stats,
               int normalize);
void release_fast_state(itnbp_net* net);
// items per launch pair chosen by the autotuner (0 before the first sweep)
int fast_chunk(itnbp_net* net, int acc_f64);
void invalidate_fast_state(itnbp_net* net);
// one-site density matrices of eligible vertices on the FP64 eigenbasis
// pipeline (rho of request i at out + offs[i]); done[i] = 1 for those served
int fast_site_rdm(itnbp_net* net, int n, const int* verts, const long long* offs, double* out,
                  std::vector<char>& done);
// half-pair environments (site, 1-based open axis) of eligible sites for
// two-site RDMs, in the generic open-contraction layout at out + offs[i]
int fast_pair_halves(itnbp_net* net, int n, const int* sites, const int* axes, const long long* offs, double* out,
                     std::vector<char>& done);

// ---- per-vertex small-chi kernel (k_vertex.hip) ----------------------------
// symmetric messages and not disabled (ITNBP_DISABLE_VERTEX=1)
bool small_path_enabled(const itnbp_net* net);
// mark the entries of seq produced by whole-vertex tasks (every chi <= 16, site
// tensor within LDS, every outgoing message in seq and not yet handled)
int small_prepare(itnbp_net* net, const std::vector<int>& seq, int acc_f64, std::vector<char>& handled);
int small_task_count(itnbp_net* net);
// batched update: the instance of every vertex and the device-side active
// flags and sweep counter the per-vertex kernels read (null pointers: off)
void small_set_batch(itnbp_net* net, const std::vector<int>& vinst, const int* active, const long long* sweep_ctr);
// times the Gram tiles of the large-site kernel on scratch outputs and keeps
// the fastest; small_big_tile reports it (0 without large sites)
int small_autotune(itnbp_net* net, int acc_f64, const void* psi, int psi_dtype, const void* msg_in, int normalize);
int small_big_tile(itnbp_net* net);
// "mfma" or "valu" for the autotuned large-site kernel ("" before tuning)
const char* small_big_kind(itnbp_net* net);
// the LDS-resident sites run as the persistent queue kernel
bool small_queue_used(itnbp_net* net);
int small_sweep(itnbp_net* net, int acc_f64, const void* psi, int psi_dtype, const void* msg_in, void* msg_out,
                MsgStat* stats, int normalize);
void release_small_state(itnbp_net* net);

// ---- bond dimensions (itnbp_api.hip) -----------------------------------------
// give every edge e the bond dimension chi_new[e] (tensors cut or zero-padded,
// messages of changed edges reset to identities); releases derived state
int net_resize_bonds(itnbp_net* net, const std::vector<int>& chi_new);

// ---- apply (k_apply.hip) ----------------------------------------------------
// resolve the device address of the no-allocation workspace on `device` (the
// current device), once per device
int apply_noalloc_init(int device);

// Runs f() and turns any C++ exception into a status code, so that none
// crosses the extern "C" boundary.
template <class F>
int itnbp_guarded(itnbp_ctx* ctx, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return itnbp_fail(ctx, ITNBP_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return itnbp_fail(ctx, ITNBP_ERR_ARG, e.what());
  } catch (...) {
    return itnbp_fail(ctx, ITNBP_ERR_ARG, "unknown error");
  }
}

#define ITNBP_RETCHK(call)          \
  do {                              \
    int _r = (call);                \
    if (_r != ITNBP_OK) return _r;  \
  } while (0)

#define ITNBP_HIPCHK(ctx, call)                                                                   \
  do {                                                                                            \
    hipError_t _e = (call);                                                                       \
    if (_e != hipSuccess)                                                                         \
      return itnbp_fail((ctx), ITNBP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)
