// Internal declarations shared by the libitnbp translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/itnbp.h"

#define ITNBP_MAXK 6                 // max degree (bonds per site tensor)
#define ITNBP_MAXR (ITNBP_MAXK + 1)  // max tensor rank (physical + bonds)
#define ITNBP_BLOCK 256

// bf16 stored as raw 16-bit patterns; conversion to wider types is exact.
struct bf16_t {
  uint16_t bits;
};

__host__ __device__ inline float bf16_to_f32(bf16_t h) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = ((uint32_t)h.bits) << 16;
  return c.f;
}

// round-to-nearest-even float -> bf16 (finite inputs)
__host__ __device__ inline bf16_t f32_to_bf16(float f) {
  union {
    uint32_t u;
    float f;
  } c;
  c.f = f;
  uint32_t lsb = (c.u >> 16) & 1u;
  uint32_t r = c.u + 0x7fffu + lsb;
  bf16_t h;
  h.bits = (uint16_t)(r >> 16);
  return h;
}

template <typename TA>
__device__ inline TA load_as(const double* p, long long i) { return (TA)p[i]; }
template <typename TA>
__device__ inline TA load_as(const float* p, long long i) { return (TA)p[i]; }
template <typename TA>
__device__ inline TA load_as(const bf16_t* p, long long i) { return (TA)bf16_to_f32(p[i]); }

// Device-side description of one site tensor and the messages flowing into it.
struct VDesc {
  long long psi_off;           // element offset of psi_v in the psi buffer
  int k;                       // number of bonds
  int D;                       // number of elements d * prod(chi)
  int dims[ITNBP_MAXR];        // dims[0] = d, dims[a] = chi of bond axis a
  long long in_msg[ITNBP_MAXR];  // offset of M_{w->v} (bond axis a >= 1) in a message buffer
};

// One "open contraction": absorb the messages on the bond axes in absorb_mask
// into psi_v, then contract with conj(psi_v) over every axis except the open
// ones (physical axis if s_open, bond axis j if j > 0).  The message update,
// the one-site RDM, half of a two-site RDM and the vertex scalar are all of
// this form.
struct OcTask {
  int v;
  int j;            // open bond axis (0: none)
  int s_open;       // physical axis open
  int absorb_mask;  // bit a: absorb incoming message on bond axis a
  long long out_off;  // output offset (elements) in the output buffer
  long long old_off;  // offset of the previous message in msg_in (-1: none)
  int normalize;
  int slot;         // statistics slot
};

// per-task statistics written by the finalisers
struct MsgStat {
  double diff;   // message_diff(new, old)
  double res2;   // ||new - old||^2 / ||new||^2
  double norm;   // ||T|| before normalisation
  double bad;    // non-finite flag
};

struct LaunchGroup {
  int kind;          // 0: small (LDS) path, 1: large (global scratch) path
  size_t begin, count;
  size_t lds;        // small: dynamic LDS bytes
  long long maxD;    // large: scratch stride
  int nsplit;        // large: split-K factor
  long long part_stride;
};

struct TaskPlan {
  std::vector<OcTask> tasks;
  std::vector<LaunchGroup> groups;
  size_t scratch_bytes = 0, partial_bytes = 0;
  int acc_f64 = 1;
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct ProfRec {
  const char* name;
  int ev0, ev1;  // indices into the event pool
};

struct itnbp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // second stream (default priority) for work that overlaps the main stream
  // inside one call, forked and joined through ev_fork / ev_join
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  std::string path;
  int num_cu = 256;
  // per-kernel timing
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<ProfRec> recs;
};

// record a timing event on the context's stream (no-op unless profiling)
int prof_mark(itnbp_ctx* ctx);
void prof_add(itnbp_ctx* ctx, const char* name, int e0, int e1);

struct itnbp_net {
  itnbp_ctx* ctx = nullptr;
  int nv = 0, ne = 0, dtype = ITNBP_F64;
  std::vector<int> edges, inc_ptr, inc_edge, chi, phys;
  std::vector<long long> psi_off, msg_off;
  long long psi_total = 0, msg_total = 0;
  std::vector<VDesc> vdesc;
  std::vector<char> msg_set;
  // every message is symmetric (M = M^T up to rounding): true after identity
  // initialisation and for sweeps from a symmetric state; the eigenbasis
  // pipeline relies on it
  bool msgs_sym = true;
  int low_x3 = 0;        // reduced-precision phase with compensated (split-operand) arithmetic
  int max_chi = 0, max_k = 0;
  // device state
  void* d_psi = nullptr;       // in dtype
  double* d_psi64 = nullptr;   // FP64 copy used by the FP64 refinement of BF16/F32 nets
  double* d_msg[2] = {nullptr, nullptr};
  float* d_msg32[2] = {nullptr, nullptr};
  int cur = 0;
  VDesc* d_vdesc = nullptr;
  DevBuf scratch, tasks, stats, partials, red, aux;
  DevBuf mdesc;          // per message: offset (long long) and chi (int), for device-side initialisation
  DevBuf apply_ws;       // itnbp_apply_gates: launch items, gates, LDS-kernel work, outputs (grows, kept)
  void* fast = nullptr;  // state of the specialised pipeline (k_fast.hip)
  void* small = nullptr; // state of the per-vertex small-chi kernel (k_vertex.hip)
};

size_t dtype_size(int dtype);

// ---- kernel launchers (k_contract.hip) ------------------------------------
// LDS bytes the small path needs for one task; 0 if it does not fit
size_t oc_small_lds_bytes(const VDesc& vd, const OcTask& t, int elem_bytes);

// Launch the small (LDS) path for tasks[0..n): every task must fit lds_bytes.
hipError_t launch_oc_small(int psi_dtype, int acc_f64, const VDesc* d_vd, const OcTask* d_tasks,
                           int n, const void* psi, const void* msg_in, void* out, MsgStat* stats,
                           size_t lds_bytes, hipStream_t s);

// Large path: global scratch; scratch must hold 2*maxD accumulator elements per
// task, partials nsplit*maxP2 per task.
hipError_t launch_oc_large(int psi_dtype, int acc_f64, const VDesc* d_vd, const OcTask* d_tasks,
                           const OcTask* h_tasks, const VDesc* h_vd, int n, const void* psi,
                           const void* msg_in, void* out, MsgStat* stats, void* scratch,
                           long long scratch_stride, void* partials, int nsplit,
                           long long part_stride, hipStream_t s);

// reduce stats[0..n) into red[0]=sum diff, red[1]=max res2, red[2]=any bad, red[3]=max diff
hipError_t launch_reduce_stats(const MsgStat* stats, int n, double* red, hipStream_t s);
// edge scalars: out[e] = sum M_{2e} .* M_{2e+1}
hipError_t launch_edge_scalars(const double* msg, const long long* offs, const int* chis, int ne,
                               double* out, hipStream_t s);
// rho2[(su,sv),(su',sv')] = sum_{b,b'} Au[(su,b),(su',b')] Av[(sv,b),(sv',b')]
hipError_t launch_rdm2_combine(const double* half, const long long* au_off, const long long* av_off,
                               const int* du, const int* dv, const int* chis, int n, double* out,
                               hipStream_t s);
hipError_t launch_convert_msgs(const double* src, float* dst, long long n, hipStream_t s);
hipError_t launch_convert_msgs_back(const float* src, double* dst, long long n, hipStream_t s);
hipError_t launch_psi_to_f64(int dtype, const void* src, double* dst, long long n, hipStream_t s);
hipError_t launch_scale_psi(int dtype, void* psi, const long long* offs, const int* sizes,
                            const double* scale, int nv, hipStream_t s);
