// Pairwise planner of small tensor-network contractions (itnbp_einsum.hpp).
#include "itnbp_einsum.hpp"

#include <algorithm>

namespace einsum_plan {

namespace {

struct Work {
  const int64_t* dim;
  int64_t max_elems;
  std::vector<std::vector<int>> lab;  // labels of every slot, in axis order
  Plan* p;
  std::string* err;

  int64_t size_of(const std::vector<int>& l) const {
    int64_t s = 1;
    for (int x : l) s *= dim[x];
    return s;
  }

  int new_slot(const std::vector<int>& l) {
    const int64_t s = size_of(l);
    if (s > max_elems) {
      *err = "an intermediate of " + std::to_string(s) + " elements exceeds the limit of " +
             std::to_string(max_elems);
      return -1;
    }
    if ((int)l.size() > kMaxRank) {
      *err = "an intermediate of rank " + std::to_string(l.size()) + " exceeds " + std::to_string(kMaxRank);
      return -1;
    }
    lab.push_back(l);
    p->sizes.push_back(s);
    return (int)lab.size() - 1;
  }

  // dst <- src with its distinct labels reordered to `keep` (every other label
  // of src summed, repeated ones traced).  Returns the new slot, or src itself
  // when src already is `keep` with no repeats.
  int unary(int src, const std::vector<int>& keep) {
    const std::vector<int>& sl = lab[src];
    if (sl == keep) return src;
    // combined stride of every distinct label of src
    std::vector<int> distinct;
    std::vector<int64_t> cstride;
    int64_t st = 1;
    for (int x : sl) {
      auto it = std::find(distinct.begin(), distinct.end(), x);
      if (it == distinct.end()) {
        distinct.push_back(x);
        cstride.push_back(st);
      } else {
        cstride[it - distinct.begin()] += st;
      }
      st *= dim[x];
    }
    Step s;
    s.u.src = src;
    for (int x : keep) {
      auto it = std::find(distinct.begin(), distinct.end(), x);
      s.u.out_dims.push_back(dim[x]);
      s.u.out_strides.push_back(cstride[it - distinct.begin()]);
    }
    for (size_t i = 0; i < distinct.size(); ++i)
      if (std::find(keep.begin(), keep.end(), distinct[i]) == keep.end()) {
        s.u.red_dims.push_back(dim[distinct[i]]);
        s.u.red_strides.push_back(cstride[i]);
      }
    if (s.u.red_dims.size() > (size_t)kMaxRank) {
      *err = "too many summed labels in one step";
      return -1;
    }
    const int d = new_slot(keep);
    if (d < 0) return -1;
    s.u.dst = d;
    p->steps.push_back(s);
    return d;
  }
};

bool has(const std::vector<int>& v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

}  // namespace

int plan(int n, const int* ranks, const int* labels, const int64_t* label_dims, int nlabels, int out_rank,
         const int* out_labels, int64_t max_elems, Plan& out, std::string& err) {
  out = Plan();
  if (n < 1 || !ranks || nlabels < 0 || out_rank < 0 || (out_rank > 0 && !out_labels) || max_elems < 1) {
    err = "malformed contraction description";
    return -1;
  }
  for (int x = 0; x < nlabels; ++x)
    if (!label_dims || label_dims[x] < 1) {
      err = "label " + std::to_string(x) + " has no positive dimension";
      return -1;
    }
  Work w{label_dims, max_elems, {}, &out, &err};
  out.ninputs = n;
  std::vector<int> outl(out_labels, out_labels + out_rank);
  int off = 0;
  for (int i = 0; i < n; ++i) {
    if (ranks[i] < 0 || ranks[i] > kMaxRank) {
      err = "tensor " + std::to_string(i) + " has rank " + std::to_string(ranks[i]) + " (at most " +
            std::to_string(kMaxRank) + ")";
      return -1;
    }
    std::vector<int> l;
    for (int a = 0; a < ranks[i]; ++a) {
      const int x = labels[off + a];
      if (x < 0 || x >= nlabels) {
        err = "tensor " + std::to_string(i) + " has a label outside [0, nlabels)";
        return -1;
      }
      l.push_back(x);
    }
    off += ranks[i];
    const int64_t s = w.size_of(l);
    if (s > max_elems) {
      err = "tensor " + std::to_string(i) + " exceeds the element limit";
      return -1;
    }
    w.lab.push_back(l);
    out.sizes.push_back(s);
  }
  for (size_t a = 0; a < outl.size(); ++a) {
    if (outl[a] < 0 || outl[a] >= nlabels) {
      err = "an output label is outside [0, nlabels)";
      return -1;
    }
    if (std::find(outl.begin() + a + 1, outl.end(), outl[a]) != outl.end()) {
      err = "output label " + std::to_string(outl[a]) + " is repeated";
      return -1;
    }
    bool found = false;
    for (int i = 0; i < n && !found; ++i) found = has(w.lab[i], outl[a]);
    if (!found) {
      err = "output label " + std::to_string(outl[a]) + " appears in no tensor";
      return -1;
    }
  }
  // a label is needed outside slot set S if it is in the output or in a live slot not in S
  std::vector<int> live;
  auto needed_outside = [&](int x, int s0, int s1) {
    if (has(outl, x)) return true;
    for (int s : live)
      if (s != s0 && s != s1 && has(w.lab[s], x)) return true;
    return false;
  };
  for (int i = 0; i < n; ++i) live.push_back(i);
  // 1. unary pre-pass: traces, labels of one tensor only, and repeated output labels
  for (size_t li = 0; li < live.size(); ++li) {
    const int s = live[li];
    std::vector<int> keep;
    for (int x : w.lab[s]) {
      if (has(keep, x)) continue;
      if (needed_outside(x, s, s)) keep.push_back(x);
    }
    const int d = w.unary(s, keep);
    if (d < 0) return -1;
    live[li] = d;
  }
  // 2. greedy pairs
  while (live.size() > 1) {
    int bi = -1, bj = -1;
    int64_t best = -1;
    bool best_shared = false;
    for (size_t i = 0; i < live.size(); ++i)
      for (size_t j = i + 1; j < live.size(); ++j) {
        const auto& A = w.lab[live[i]];
        const auto& B = w.lab[live[j]];
        bool shared = false;
        for (int x : A) shared = shared || has(B, x);
        // result labels: union minus the shared labels needed nowhere else
        std::vector<int> res;
        for (int x : A)
          if (!has(B, x) || needed_outside(x, live[i], live[j])) res.push_back(x);
        for (int x : B)
          if (!has(A, x)) res.push_back(x);
        const int64_t rs = w.size_of(res);
        if (bi < 0 || (shared && !best_shared) || (shared == best_shared && rs < best)) {
          bi = (int)i;
          bj = (int)j;
          best = rs;
          best_shared = shared;
        }
      }
    const int sa = live[bi], sb = live[bj];
    const std::vector<int> A = w.lab[sa], B = w.lab[sb];
    std::vector<int> m, nn, k, b;
    for (int x : A) {
      if (!has(B, x))
        m.push_back(x);
      else if (needed_outside(x, sa, sb))
        b.push_back(x);
      else
        k.push_back(x);
    }
    for (int x : B)
      if (!has(A, x)) nn.push_back(x);
    std::vector<int> ao = m, bo = nn, co = m;
    ao.insert(ao.end(), k.begin(), k.end());
    ao.insert(ao.end(), b.begin(), b.end());
    bo.insert(bo.end(), k.begin(), k.end());
    bo.insert(bo.end(), b.begin(), b.end());
    co.insert(co.end(), nn.begin(), nn.end());
    co.insert(co.end(), b.begin(), b.end());
    const int pa = w.unary(sa, ao);
    if (pa < 0) return -1;
    const int pb = w.unary(sb, bo);
    if (pb < 0) return -1;
    const int pc = w.new_slot(co);
    if (pc < 0) return -1;
    Step s;
    s.gemm = true;
    s.g.a = pa;
    s.g.bt = pb;
    s.g.dst = pc;
    s.g.m = w.size_of(m);
    s.g.n = w.size_of(nn);
    s.g.k = w.size_of(k);
    s.g.batch = w.size_of(b);
    out.flops += s.g.m * s.g.n * s.g.k * s.g.batch;
    out.steps.push_back(s);
    live.erase(live.begin() + bj);
    live.erase(live.begin() + bi);
    live.push_back(pc);
  }
  // 3. output order
  const int last = live[0];
  for (int x : w.lab[last])
    if (!has(outl, x)) {
      err = "internal: label " + std::to_string(x) + " survived the contraction";
      return -1;
    }
  const int r = w.unary(last, outl);
  if (r < 0) return -1;
  // the result must be a slot of its own: an input passed through unchanged is copied
  if (r < n) {
    Step s;
    s.u.src = r;
    int64_t st = 1;
    for (int x : outl) {
      s.u.out_dims.push_back(label_dims[x]);
      s.u.out_strides.push_back(st);
      st *= label_dims[x];
    }
    const int d = w.new_slot(outl);
    if (d < 0) return -1;
    s.u.dst = d;
    out.steps.push_back(s);
    out.result = d;
  } else {
    out.result = r;
  }
  return 0;
}

}  // namespace einsum_plan
