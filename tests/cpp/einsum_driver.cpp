// The contraction planner (itnbp_einsum.cpp) driven on its own, for the
// sanitizer build of the tests: random label sets (traces, summed labels,
// hyperedges, outer products, scalars, permutations) are planned, the plan is
// executed by a CPU restatement of the device steps (unary: permute + trace +
// sum; GEMM: C[m, n, b] = sum_k A[m, k, b] Bt[n, k, b]), and the result is
// compared with a brute-force sum over every assignment of the labels.
// Prints {"cases": N, "checks": M, "failures": F, "max_err": E} as JSON.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../itensornetworks.jl_amd/csrc/itnbp_einsum.hpp"

using namespace einsum_plan;

namespace {

int failures = 0;
long checks = 0;
double max_err = 0;

void expect(bool ok, const std::string& what) {
  ++checks;
  if (!ok) {
    ++failures;
    fprintf(stderr, "FAILED: %s\n", what.c_str());
  }
}

std::vector<double> execute(const Plan& p, const std::vector<std::vector<double>>& inputs) {
  std::vector<std::vector<double>> slot(p.sizes.size());
  for (int i = 0; i < p.ninputs; ++i) slot[i] = inputs[i];
  for (const Step& s : p.steps) {
    if (s.gemm) {
      const Gemm& g = s.g;
      const auto& A = slot[g.a];
      const auto& B = slot[g.bt];
      std::vector<double> C((size_t)(g.m * g.n * g.batch), 0.0);
      expect((int64_t)A.size() == g.m * g.k * g.batch && (int64_t)B.size() == g.n * g.k * g.batch,
             "GEMM operand sizes");
      for (int64_t b = 0; b < g.batch; ++b)
        for (int64_t n = 0; n < g.n; ++n)
          for (int64_t m = 0; m < g.m; ++m) {
            double acc = 0;
            for (int64_t k = 0; k < g.k; ++k) acc += A[m + g.m * (k + g.k * b)] * B[n + g.n * (k + g.k * b)];
            C[m + g.m * (n + g.n * b)] = acc;
          }
      expect((int64_t)C.size() == p.sizes[g.dst], "GEMM result size");
      slot[g.dst] = C;
    } else {
      const Unary& u = s.u;
      const auto& S = slot[u.src];
      int64_t ot = 1, rt = 1;
      for (int64_t d : u.out_dims) ot *= d;
      for (int64_t d : u.red_dims) rt *= d;
      std::vector<double> D((size_t)ot);
      for (int64_t o = 0; o < ot; ++o) {
        int64_t rem = o, base = 0;
        for (size_t i = 0; i < u.out_dims.size(); ++i) {
          base += (rem % u.out_dims[i]) * u.out_strides[i];
          rem /= u.out_dims[i];
        }
        double acc = 0;
        for (int64_t r = 0; r < rt; ++r) {
          int64_t rr = r, off = base;
          for (size_t i = 0; i < u.red_dims.size(); ++i) {
            off += (rr % u.red_dims[i]) * u.red_strides[i];
            rr /= u.red_dims[i];
          }
          expect(off >= 0 && off < (int64_t)S.size(), "unary read in bounds");
          if (off >= 0 && off < (int64_t)S.size()) acc += S[off];
        }
        D[o] = acc;
      }
      expect(ot == p.sizes[u.dst], "unary result size");
      slot[u.dst] = D;
    }
  }
  return slot[p.result];
}

std::vector<double> brute(int n, const std::vector<std::vector<int>>& lab, const std::vector<int64_t>& dim,
                          const std::vector<std::vector<double>>& in, const std::vector<int>& outl) {
  const int nl = (int)dim.size();
  // labels in no tensor take part in no sum
  std::vector<int64_t> run(nl, 1);
  for (int i = 0; i < n; ++i)
    for (int x : lab[i]) run[x] = dim[x];
  int64_t total = 1, osz = 1;
  for (int64_t d : run) total *= d;
  for (int x : outl) osz *= dim[x];
  std::vector<double> out((size_t)osz, 0.0);
  std::vector<int64_t> val(nl);
  for (int64_t a = 0; a < total; ++a) {
    int64_t r = a;
    for (int x = 0; x < nl; ++x) {
      val[x] = r % run[x];
      r /= run[x];
    }
    double prod = 1;
    for (int i = 0; i < n; ++i) {
      int64_t off = 0, st = 1;
      for (int x : lab[i]) {
        off += val[x] * st;
        st *= dim[x];
      }
      prod *= in[i][off];
    }
    int64_t o = 0, st = 1;
    for (int x : outl) {
      o += val[x] * st;
      st *= dim[x];
    }
    out[o] += prod;
  }
  return out;
}

void one_case(std::mt19937& rng, int c) {
  std::uniform_int_distribution<int> U(0, 1 << 30);
  const int nl = 1 + U(rng) % 6;
  std::vector<int64_t> dim(nl);
  for (auto& d : dim) d = 1 + U(rng) % 3;
  const int n = 1 + U(rng) % 4;
  std::vector<std::vector<int>> lab(n);
  std::vector<int> ranks, flat;
  for (int i = 0; i < n; ++i) {
    const int r = U(rng) % 5;
    for (int a = 0; a < r; ++a) lab[i].push_back(U(rng) % nl);
    ranks.push_back(r);
    flat.insert(flat.end(), lab[i].begin(), lab[i].end());
  }
  // output: a random subset of the labels present, each once, shuffled
  std::vector<int> present;
  for (int i = 0; i < n; ++i)
    for (int x : lab[i])
      if (std::find(present.begin(), present.end(), x) == present.end()) present.push_back(x);
  std::vector<int> outl;
  for (int x : present)
    if (U(rng) % 2) outl.push_back(x);
  std::shuffle(outl.begin(), outl.end(), rng);
  std::vector<std::vector<double>> in(n);
  std::normal_distribution<double> N01;
  for (int i = 0; i < n; ++i) {
    int64_t s = 1;
    for (int x : lab[i]) s *= dim[x];
    in[i].resize((size_t)s);
    for (auto& v : in[i]) v = N01(rng);
  }
  Plan p;
  std::string err;
  const int rc = plan(n, ranks.data(), flat.data(), dim.data(), nl, (int)outl.size(), outl.data(), 1 << 20, p, err);
  expect(rc == 0, "case " + std::to_string(c) + " plans: " + err);
  if (rc != 0) return;
  expect(p.result >= p.ninputs, "the result is a slot of its own");
  const auto got = execute(p, in);
  const auto want = brute(n, lab, dim, in, outl);
  expect(got.size() == want.size(), "result size");
  for (size_t i = 0; i < got.size() && i < want.size(); ++i) {
    const double e = std::fabs(got[i] - want[i]) / (1.0 + std::fabs(want[i]));
    max_err = std::max(max_err, e);
    expect(e < 1e-12, "case " + std::to_string(c) + " element " + std::to_string(i));
  }
}

void invalid_cases() {
  Plan p;
  std::string err;
  const int64_t dims[2] = {2, 3};
  const int ranks[1] = {2};
  const int labels[2] = {0, 1};
  const int out_missing[1] = {1};
  const int out_rep[2] = {0, 0};
  const int bad_label[2] = {0, 5};
  expect(plan(1, ranks, labels, dims, 2, 2, out_rep, 1 << 20, p, err) != 0, "repeated output label rejected");
  expect(plan(1, ranks, bad_label, dims, 2, 1, out_missing, 1 << 20, p, err) != 0, "label out of range rejected");
  expect(plan(1, ranks, labels, dims, 2, 1, out_missing, 5, p, err) != 0, "element limit enforced");
  const int64_t zero[2] = {0, 3};
  expect(plan(1, ranks, labels, zero, 2, 1, out_missing, 1 << 20, p, err) != 0, "zero dimension rejected");
  const int ranks2[2] = {2, 1};
  const int labels2[3] = {0, 1, 0};
  const int out3[1] = {2};
  const int64_t dims3[3] = {2, 3, 4};
  expect(plan(2, ranks2, labels2, dims3, 3, 1, out3, 1 << 20, p, err) != 0, "output label in no tensor rejected");
}

}  // namespace

int main() {
  std::mt19937 rng(12345);
  const int ncases = 400;
  for (int c = 0; c < ncases; ++c) one_case(rng, c);
  invalid_cases();
  printf("{\"cases\": %d, \"checks\": %ld, \"failures\": %d, \"max_err\": %.3e}\n", ncases, checks, failures,
         max_err);
  return failures == 0 ? 0 : 1;
}
