// Pairwise plan of a small real tensor-network contraction (an einsum over
// integer labels), free of HIP so that it builds with a plain C++ compiler and
// runs under the sanitizers in the tests.  The device executor (k_einsum.hip)
// runs the plan's steps as they are; the planner does all index bookkeeping.
//
// Semantics: tensor i has rank[i] axes, column-major (first axis fastest),
// labelled from [0, nlabels); every label has one dimension.  A label repeated
// within one tensor is traced (its diagonal), a label that appears in no
// other tensor and not in the output is summed, a label shared by several
// tensors and absent from the output is contracted, and one in the output is
// kept (a label shared by more than two tensors is carried as a batch index
// until its last pair).  The output axes are out_labels, in that order; each
// must appear in some input, at most once per tensor.
//
// Plan: unary steps (permute + trace + sum) on the inputs that need them, then
// greedy pairwise products (the pair whose result is smallest first; a pair
// sharing no label only when no pair shares one), each a batched GEMM
//   C[m, n, b] = sum_k A[m, k, b] Bt[n, k, b]
// on operands permuted by unary steps into those orders, then a final
// permutation into the output order.  This stands in for the contraction
// sequence the reference finds for a partition's factors
// (src/contraction_sequences.jl:9-20) before contracting them
// (src/caches/abstractbeliefpropagationcache.jl:225-239).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace einsum_plan {

constexpr int kMaxRank = 16;  // axes of any tensor a step reads or writes

// out[o] = sum over r of src[sum_a o_a out_strides[a] + sum_c r_c red_strides[c]],
// o and r column-major over out_dims and red_dims
struct Unary {
  int src = -1, dst = -1;
  std::vector<int64_t> out_dims, out_strides;
  std::vector<int64_t> red_dims, red_strides;
};

// C[m + M (n + N b)] = sum_k A[m + M (k + K b)] Bt[n + N (k + K b)]
struct Gemm {
  int a = -1, bt = -1, dst = -1;
  int64_t m = 1, n = 1, k = 1, batch = 1;
};

struct Step {
  bool gemm = false;
  Unary u;
  Gemm g;
};

struct Plan {
  int ninputs = 0;
  std::vector<int64_t> sizes;  // elements of every slot; slots [0, ninputs) are the inputs
  std::vector<Step> steps;
  int result = -1;             // slot holding the output (in out_labels order)
  int64_t flops = 0;           // multiply-adds of the GEMM steps
};

// Plans the contraction.  Returns 0, or -1 with the reason in err (malformed
// description, an output label missing from the inputs or repeated, a tensor
// or intermediate above max_elems elements, a rank above kMaxRank).
int plan(int n, const int* ranks, const int* labels, const int64_t* label_dims, int nlabels, int out_rank,
         const int* out_labels, int64_t max_elems, Plan& out, std::string& err);

}  // namespace einsum_plan
