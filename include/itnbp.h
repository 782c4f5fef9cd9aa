/*
 * itnbp.h - C ABI of libitnbp.so, the MI355X belief-propagation library.
 *
 * Drop-in boundary for the BP hot path of ITensorNetworks.jl: the Julia shim
 * (ccall) or the Python host package (ctypes) marshals an ITensorNetwork's
 * site tensors into flat column-major buffers and calls these entry points;
 * every array crossing the boundary is a plain pointer plus sizes, owned by
 * the caller, and copied by the library (no caller pointer is kept after a
 * call returns).
 *
 * Conventions
 *   - Network: nv vertices, ne undirected edges edges[2e], edges[2e+1].
 *     Directed message ids: 2e is edges[2e] -> edges[2e+1], 2e+1 the reverse.
 *   - Bond (axis) order per vertex: CSR inc_ptr[nv+1] / inc_edge[2ne]; axis
 *     a+1 of psi_v is the bond inc_edge[inc_ptr[v]+a]; axis 0 is physical.
 *     (test/utils.jl:32-36: inds_v = [siteinds[v]; link_v])
 *   - Site tensors: column-major (first index fastest), [s, b_1, ..., b_k],
 *     concatenated over vertices in vertex order.
 *   - Messages: chi x chi column-major M[b_ket, b_bra]
 *     (formnetworks/quadraticformnetwork.jl:115-119), concatenated by id.
 *   - One-site operators: d x d column-major O[s', s] (ITensors op(name, s));
 *     two-site: (d_u d_v) x (d_u d_v) O[(s_u', s_v'), (s_u, s_v)], s_u fastest.
 *   - Every call returns ITNBP_OK or an error code; itnbp_last_error() gives
 *     the text. No C++ exception crosses the ABI. Not thread-safe per context.
 */
#ifndef ITNBP_H
#define ITNBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ITNBP_OK = 0,
  ITNBP_ERR_ARG = 1,         /* invalid argument / shape */
  ITNBP_ERR_HIP = 2,         /* HIP runtime error */
  ITNBP_ERR_NOMEM = 3,       /* device allocation failed */
  ITNBP_ERR_NONFINITE = 4,   /* NaN/Inf detected in a result */
  ITNBP_ERR_UNSUPPORTED = 5, /* shape or mode this build does not handle */
  ITNBP_ERR_STATE = 6        /* e.g. message read before it was set */
};

/* storage element types of site tensors */
enum { ITNBP_F64 = 0, ITNBP_F32 = 1, ITNBP_BF16 = 2 };

/* BP schedules (edge_sequences.jl:9-51) */
enum { ITNBP_SCHED_SEQUENTIAL = 0, ITNBP_SCHED_PARALLEL = 1 };

/* arithmetic of the message contractions */
enum {
  ITNBP_PREC_F64 = 0,         /* FP64 everywhere (the plain path) */
  ITNBP_PREC_F32 = 1,         /* FP32 messages; FP32 arithmetic, except that with BF16
                                 storage the chi = 32 pipelines run 16-bit MFMA
                                 (FP16 intermediates, FP32 accumulation) */
  ITNBP_PREC_BF16_F64 = 2,    /* BF16 MFMA with FP32 accumulation until the
                                 residual < switch_tol, then FP64 refinement */
  ITNBP_PREC_BF16X3 = 3,      /* compensated BF16 MFMA: every rounded operand split
                                 into three BF16 terms (about 24 significant bits,
                                 FP32-class results), FP32 intermediates and
                                 messages; where the chi = 32 eigenbasis pipeline
                                 does not apply, the absorbed-pair passes with
                                 two-term splits or the FP32 kernels */
  ITNBP_PREC_BF16X3_F64 = 4,  /* ITNBP_PREC_BF16X3 until the residual < switch_tol,
                                 then FP64 refinement */
  ITNBP_PREC_BF16_BF16X3 = 5, /* progressive: the ITNBP_PREC_F32 kernels until the
                                 residual < switch_tol (or stops decreasing), then
                                 ITNBP_PREC_BF16X3 */
  ITNBP_PREC_BF16_BF16X3_F64 = 6, /* the ITNBP_PREC_F32 kernels until the residual
                                 < 1e-2, ITNBP_PREC_BF16X3 until < switch_tol (or
                                 stagnating), then FP64 refinement */
  ITNBP_PREC_F16X2 = 7,       /* compensated FP16 MFMA on the chi = 32 eigenbasis
                                 pipeline: every rounded operand split into two FP16
                                 terms (about 22 significant bits), FP32 intermediates
                                 and messages; elsewhere as ITNBP_PREC_BF16X3 */
  ITNBP_PREC_BF16_F16X2 = 8   /* progressive: the ITNBP_PREC_F32 kernels until the
                                 residual < switch_tol (or stops decreasing), then
                                 ITNBP_PREC_F16X2 */
};

typedef struct itnbp_ctx itnbp_ctx;
typedef struct itnbp_net itnbp_net;

typedef struct {
  int schedule;          /* ITNBP_SCHED_* */
  int maxiter;           /* >= 1; required (abstractbeliefpropagationcache.jl:315-317) */
  double tol;            /* stop when mean message_diff <= tol; < 0 disables */
  double res_tol;        /* stop when max ||new-old||/||new|| <= res_tol; < 0 disables */
  int precision;         /* ITNBP_PREC_* */
  double switch_tol;     /* BF16 phase stops when max residual < switch_tol */
  int normalize;         /* divide each message by its Frobenius norm (default 1) */
  int nseq;              /* length of seq (0: default schedule of the network) */
  const int* seq;        /* directed message ids in update order */
  int verbose;
} itnbp_update_opts;

typedef struct {
  int iterations;        /* sweeps run (all phases) */
  int iterations_low;    /* sweeps of the reduced-precision phase */
  double mean_diff;      /* mean message_diff of the last sweep */
  double max_residual;   /* max ||new-old||_F/||new||_F of the last sweep */
  int converged;
  double device_ms;      /* device time of the sweeps (HIP events) */
} itnbp_update_result;

/* ---- context ---------------------------------------------------------- */
int itnbp_create(int device, itnbp_ctx** out);
int itnbp_destroy(itnbp_ctx* ctx);
const char* itnbp_last_error(itnbp_ctx* ctx);
int itnbp_synchronize(itnbp_ctx* ctx);

/* ---- network ---------------------------------------------------------- */
/* Replaces BeliefPropagationCache(QuadraticFormNetwork(psi)) construction
 * (caches/beliefpropagationcache.jl:20-35, formnetworks/quadraticformnetwork.jl:75-89):
 * the bra layer is conj(psi) on the fly and the identity operator is elided. */
int itnbp_net_create(itnbp_ctx* ctx, int nv, int ne, const int* edges,
                     const int* inc_ptr, const int* inc_edge, const int* bond_dims,
                     const int* phys_dims, int dtype, itnbp_net** out);
int itnbp_net_destroy(itnbp_net* net);
/* deep copy (functional API: update/set_message return new caches,
 * caches/beliefpropagationcache.jl:43-47) */
int itnbp_net_copy(itnbp_net* src, itnbp_net** out);
int64_t itnbp_net_psi_numel(itnbp_net* net);
/* current bond dimension of every edge (they change with itnbp_apply_gates);
 * linkinds / dim of the reference network (src/abstractitensornetwork.jl) */
int itnbp_get_bond_dims(itnbp_net* net, int* bond_dims);
int64_t itnbp_net_msg_numel(itnbp_net* net);

/* site tensors, all vertices, column-major, in the net's dtype
 * (bpc[v] = ... of caches/abstractbeliefpropagationcache.jl:158-171) */
int itnbp_set_sites(itnbp_net* net, const void* data);
int itnbp_get_sites(itnbp_net* net, void* data);
int itnbp_set_site(itnbp_net* net, int v, const void* data);

/* messages (FP64 host buffers), all directed ids
 * (messages / set_messages!, abstractbeliefpropagationcache.jl:173-200) */
int itnbp_set_messages(itnbp_net* net, const double* data);
int itnbp_get_messages(itnbp_net* net, double* data);
int itnbp_set_message(itnbp_net* net, int msg_id, const double* data);
int itnbp_get_message(itnbp_net* net, int msg_id, double* data);
/* identity_messages (formnetworks/quadraticformnetwork.jl:96-124) */
int itnbp_init_identity_messages(itnbp_net* net);
/* 1 if message id has been set (message() has no default,
 * abstractbeliefpropagationcache.jl:173-175) */
int itnbp_message_is_set(itnbp_net* net, int msg_id);

/* default sequential schedule (edge_sequences.jl:32-47); out has 2*ne ids */
int itnbp_forest_cover_sequence(itnbp_net* net, int* out);

/* ---- BP --------------------------------------------------------------- */
/* update(bpc; alg="bp", maxiter, tol, edge_sequence, message_update_alg)
 * (abstractbeliefpropagationcache.jl:313-337) */
int itnbp_update(itnbp_net* net, const itnbp_update_opts* opts, itnbp_update_result* res);
/* updated_message(Algorithm"contract", bpc, e) for a list of directed ids
 * (abstractbeliefpropagationcache.jl:225-239); out: chi^2 per id, FP64 */
int itnbp_updated_messages(itnbp_net* net, int n, const int* msg_ids, int normalize,
                           double* out);

/* ---- environments / expectation values / scalars ---------------------- */
/* With every message symmetric, degree-4 and degree-3 sites with d = 2 and every
 * chi = 32 take the eigenbasis route (FP64 rotations of the site into the
 * eigenbasis of its incoming messages, then weighted Grams) when a request holds
 * at least ITNBP_RDM_FAST_MIN (default 16) such sites or half-pairs; the others
 * use the generic open contraction.  Both are FP64 and agree to rounding. */
/* one-site RDMs rho_v[s, s'] (d x d, unnormalised) of the BP environment
 * (environment(bpc, [operator vertex]), caches/beliefpropagationcache.jl:100-105;
 * expect.jl:5-19 contracts it with op and with delta) */
int itnbp_rdm1(itnbp_net* net, int n, const int* verts, double* out);
/* <O_v> = tr(O rho_v)/tr(rho_v); ops: n d x d matrices (expect.jl:5-41) */
int itnbp_expect1(itnbp_net* net, int n, const int* verts, const double* ops, double* out);
/* two-site RDMs of neighbouring pairs given by edge ids, (d_u d_v)^2 each
 * (test/test_belief_propagation.jl:64-91 pattern) */
int itnbp_rdm2(itnbp_net* net, int n, const int* edge_ids, double* out);
int itnbp_expect2(itnbp_net* net, int n, const int* edge_ids, const double* ops, double* out);
/* region_scalar for every vertex and every edge (caches/beliefpropagationcache.jl:107-119) */
int itnbp_region_scalars(itnbp_net* net, double* vertex_scalars, double* edge_scalars);
/* logscalar(bpc) (abstractbeliefpropagationcache.jl:397-408): re/im parts;
 * *is_neg_inf set when an edge scalar is 0 */
int itnbp_logscalar(itnbp_net* net, double* re, double* im, int* is_neg_inf);
/* rescale_messages + rescale_partitions as used by normalize(alg="bp")
 * (caches/beliefpropagationcache.jl:121-139, abstractbeliefpropagationcache.jl:349-395,
 * normalize.jl:125-142) */
int itnbp_normalize(itnbp_net* net);

/* ---- BP-gauged two-site gates ----------------------------------------- */
/* apply(o, psi; envs, maxdim, cutoff, normalize, callback) for n gates on
 * disjoint bonds, envs = the net's current messages (apply.jl:33-146).
 * gates: n (d_u d_v)^2 FP64 matrices. maxdim <= 0: no cap; cutoff < 0: none.
 * truncerr[n]; svals[n*max_keep] (zero padded) may be NULL; kept[n] may be NULL.
 * Each gate's bond takes the kept dimension (apply.jl:81-93): it grows or
 * shrinks on the device; the messages of a bond whose dimension changed become
 * identities of the new dimension; itnbp_get_bond_dims reports the new
 * dimensions.  A gate whose two-site tensor is not finite leaves its bond
 * untouched (kept = -1) and the call returns ITNBP_ERR_NONFINITE after the
 * other gates were applied.  Limits: degree <= 4, external bonds <= 32,
 * d chi <= 64 on the gate's bond, Theta <= 128 x 128, kept rank <= 64
 * (bonds with external bonds <= 8 and Theta <= 32 x 32 run on the LDS
 * kernel, the others on its global-memory variant). */
int itnbp_apply_gates(itnbp_net* net, int n, const int* edge_ids, const double* gates,
                      int maxdim, double cutoff, int normalize, double* truncerr,
                      double* svals, int max_keep, int* kept);
/* one-site gates psi_v <- O psi_v on n distinct vertices (apply.jl:110-116):
 * ops: n d x d matrices O[s', s]; normalize divides each updated tensor by its
 * Frobenius norm (:113-115); norms[n] (may be NULL) receives ||O psi_v||. */
int itnbp_apply_site_gates(itnbp_net* net, int n, const int* verts, const double* ops, int normalize,
                           double* norms);
/* itnbp_apply_gates without any device allocation: the launch items, gate
 * matrices, per-gate workspace and per-gate outputs live in one file-scope
 * device array of the library (18.4 MiB per GPU, held until the process
 * ends).  No counterpart in the reference: for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops).  Limits:
 * n <= 128 gates, (d_u d_v)^2 <= 256 elements per gate, max_keep <= 32;
 * larger calls return ITNBP_ERR_UNSUPPORTED.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other threads, other
 * contexts on the same device) race. */
int itnbp_apply_gates_noalloc(itnbp_net* net, int n, const int* edge_ids, const double* gates,
                              int maxdim, double cutoff, int normalize, double* truncerr,
                              double* svals, int max_keep, int* kept);

/* ---- adjoint ---------------------------------------------------------- */
/* BP energy E = sum_e <O2_e> + sum_v <O1_v> after nsweeps Jacobi sweeps from
 * identity messages, and dE/dpsi for every site tensor (FP64), by hand-written
 * adjoint kernels; sweeps are checkpointed every `checkpoint` sweeps (<= 0:
 * every sweep) and recomputed in the backward pass. No counterpart in the
 * reference (it has no autodiff); the Julia shim wraps it as a ChainRulesCore rrule.
 * ops1: nv d x d (may be NULL), ops2: ne (d_u d_v)^2 (may be NULL). */
int itnbp_energy_grad(itnbp_net* net, int nsweeps, int checkpoint, const double* ops1,
                      const double* ops2, double* energy, double* grad);
/* Pullback through nsweeps Jacobi sweeps from identity messages (the forward of
 * itnbp_energy_grad): given the cotangent mbar of the final messages (all
 * messages, itnbp_net_msg_numel doubles), the gradient with respect to every
 * site tensor (itnbp_net_psi_numel doubles). With itnbp_apply_vjp it gives the
 * gradient of a TEBD step (BP sweeps, then a gate round). No counterpart in
 * the reference. Leaves the final forward messages in the network. */
int itnbp_sweeps_vjp(itnbp_net* net, int nsweeps, int checkpoint, const double* mbar, double* grad);
/* Vector-Jacobian product of itnbp_apply_gates (simple_update_bp, apply.jl:33-95,
 * and the normalisation of :134-137) for n gates on disjoint bonds; no
 * counterpart in the reference (it has no autodiff). Called on the network as
 * it was before the gates, with the forward call's arguments and the ranks it
 * kept. dpsi_out: per gate the cotangents of psi_u' then psi_v' (u, v = the
 * edge's endpoints in the order of itnbp_net_create; column-major, the bond of
 * dimension kept[i]). Writes dpsi_in (per gate psi_u then psi_v, the current
 * shapes) and dgates (the gates' layout), and adds the cotangents of the
 * environment messages into dmsgs (all messages, itnbp_net_msg_numel
 * doubles). ITNBP_ERR_ARG if kept differs from the recomputed rank.
 * Limits: those of itnbp_apply_gates' LDS kernel (external bonds <= 8, Theta
 * <= 32 x 32); other bonds return ITNBP_ERR_UNSUPPORTED. */
int itnbp_apply_vjp(itnbp_net* net, int n, const int* edge_ids, const double* gates, int maxdim,
                    double cutoff, int normalize, const int* kept, const double* dpsi_out,
                    double* dpsi_in, double* dgates, double* dmsgs);

/* ---- batched heterogeneous BP ----------------------------------------- */
/* n networks (any lattices, bond dims, same dtype) updated together: every
 * message of every network in one launch per sweep; converged networks drop
 * out (Jacobi schedule). res may hold n results. */
int itnbp_update_batched(int n, itnbp_net** nets, const itnbp_update_opts* opts,
                         itnbp_update_result* res);

/* ---- two-layer networks: bra != ket, real or complex ------------------ */
/* BP on <phi|psi>, the LinearFormNetwork of inner(phi, psi; alg="bp")
 * (inner.jl:128-160, formnetworks/linearformnetwork.jl:15-25), one {bra, ket}
 * partition per site; with the operator folded into the ket, on <phi|A|psi>
 * (inner(phi, A, psi; alg="bp"), formnetworks/bilinearformnetwork.jl:23-42);
 * with phi = psi, the norm network of a complex state. FP64 or complex FP64
 * (is_complex: every element an interleaved (re, im) pair of doubles).
 * Site tensors column-major [s, bonds in incident order]; the bra layer holds
 * phi itself (the conjugation is done by the kernels). Messages
 * M[b_ket, b_bra] (chi_ket x chi_bra column-major); no default message
 * (initialize_cache.jl:10-12 builds the cache without one). */
typedef struct itnbp_lnet itnbp_lnet;
int itnbp_lnet_create(itnbp_ctx* ctx, int nv, int ne, const int* edges, const int* inc_ptr,
                      const int* inc_edge, const int* ket_bond_dims, const int* bra_bond_dims,
                      const int* phys_dims, int is_complex, itnbp_lnet** out);
int itnbp_lnet_destroy(itnbp_lnet* lnet);
int64_t itnbp_lnet_msg_numel(itnbp_lnet* lnet);
int64_t itnbp_lnet_ket_numel(itnbp_lnet* lnet);
int64_t itnbp_lnet_bra_numel(itnbp_lnet* lnet);
int itnbp_lnet_set_tensors(itnbp_lnet* lnet, const double* ket, const double* bra);
/* ket_v <- A_v psi_v: ops are A_v[s', s, a_1 .. a_k] (s' the bra side), the
 * merged bond index is b + chi_psi a, and the network's ket bond dimensions
 * must be chi_psi * chi_op (bilinearformnetwork.jl:23-42 with the operator
 * and ket link indices of every edge merged) */
int itnbp_lnet_fold_operator(itnbp_lnet* lnet, const double* psi, const int* psi_bond_dims,
                             const double* ops, const int* op_bond_dims);
int itnbp_lnet_get_ket(itnbp_lnet* lnet, double* ket);
int itnbp_lnet_set_message(itnbp_lnet* lnet, int msg_id, const double* data);
int itnbp_lnet_get_message(itnbp_lnet* lnet, int msg_id, double* data);
int itnbp_lnet_message_is_set(itnbp_lnet* lnet, int msg_id);
/* delta(b_ket, b_bra) on every message (equal bond dimensions only) */
int itnbp_lnet_init_identity_messages(itnbp_lnet* lnet);
/* update(bpc; alg="bp") (abstractbeliefpropagationcache.jl:294-337); FP64 only */
int itnbp_lnet_update(itnbp_lnet* lnet, const itnbp_update_opts* opts, itnbp_update_result* res);
/* vertex scalars (nv (re, im) pairs) and edge scalars (ne pairs)
 * (beliefpropagationcache.jl:107-119); logscalar: sum of the complex logs of
 * the vertex scalars minus those of the edge scalars (:397-412) */
int itnbp_lnet_region_scalars(itnbp_lnet* lnet, double* vertex_scalars, double* edge_scalars);
int itnbp_lnet_logscalar(itnbp_lnet* lnet, double* re, double* im);
/* R_v[s_ket, s_bra] = environment of v contracted with ket and conj(bra) over
 * every bond (beliefpropagationcache.jl:100-105), unnormalised, d x d each */
int itnbp_lnet_rdm1(itnbp_lnet* lnet, int n, const int* verts, double* out);

/* ---- small tensor-network contractions ------------------------------- */
/* Contraction of n real FP64 tensors on the device, for BP over generalised
 * partitions: the tensor of a partition is the contraction of its sites'
 * tensors over the bonds inside it (the factors the reference contracts with a
 * partition's incoming messages, src/caches/abstractbeliefpropagationcache.jl:
 * 225-239, for the partitioned cache of beliefpropagationcache.jl:20-35), and
 * a member site's density matrix is a partial trace of its partition's.
 * Tensor i has ranks[i] axes, column-major, labelled by the next ranks[i]
 * entries of labels (ids in [0, nlabels), dimension label_dims[id]); data holds
 * the tensors back to back.  A label repeated within one tensor is traced, one
 * in no other tensor and not in the output is summed, one shared and not in
 * the output is contracted.  out (column-major over out_labels, each at most
 * once) receives the result.  Pairwise order: the smallest intermediate first
 * (standing in for contraction_sequence, src/contraction_sequences.jl:9-20);
 * every product runs as a batched GEMM on the FP64 matrix cores.  Limits: rank
 * 16 per tensor, 2^28 elements per tensor and 2^30 per call
 * (ITNBP_ERR_UNSUPPORTED beyond). */
int itnbp_einsum(itnbp_ctx* ctx, int n, const int* ranks, const int* labels, const int* label_dims, int nlabels,
                 const double* data, int out_rank, const int* out_labels, double* out);

/* ---- diagnostics ------------------------------------------------------ */
/* name of the kernel path the last itnbp_update used */
const char* itnbp_last_path(itnbp_ctx* ctx);
/* per-kernel timing with HIP events recorded on the library's stream around
 * every launch while enabled; query returns the launch count and the summed
 * duration (ms) of the kernel called `name` since the last reset. No
 * counterpart in the reference (it has no timers); used by the benchmark. */
int itnbp_profile_enable(itnbp_ctx* ctx, int enable);
int itnbp_profile_reset(itnbp_ctx* ctx);
int itnbp_profile_query(itnbp_ctx* ctx, const char* name, int* count, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
