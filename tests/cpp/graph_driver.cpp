// Host-side topology logic (itnbp_graph.cpp) driven on its own, for the
// sanitizer build of the tests (g++ -fsanitize=address,undefined): no GPU, no
// HIP.
//
//   graph_driver <graph.bin>   checks the invariants on one network graph and
//                              prints its sequences as JSON
//   graph_driver --invalid     runs the malformed descriptions and prints the
//                              error codes as JSON
//
// graph.bin (little-endian int32): nv, ne, edges[2 ne], inc_ptr[nv + 1],
// inc_edge[2 ne], chi[ne], phys[nv].
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../itensornetworks.jl_amd/csrc/itnbp_graph.hpp"
#include "../../include/itnbp.h"

using namespace itnbp_graph;

namespace {

int failures = 0;
long checks = 0;

void expect(bool ok, const std::string& what) {
  ++checks;
  if (!ok) {
    ++failures;
    fprintf(stderr, "FAILED: %s\n", what.c_str());
  }
}

bool read_all(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

const Limits kLim{6, 64, 8, 1LL << 30};

// messages read by the update of m: those into its source from every other neighbour
std::vector<int> inputs_of(const Topo& t, int m) {
  std::vector<int> in;
  const int u = msg_src(t, m), v = msg_dst(t, m);
  for (int a = t.inc_ptr[u]; a < t.inc_ptr[u + 1]; ++a) {
    const int x = msg_id_into(t, t.inc_edge[a], u);
    if (msg_src(t, x) != v) in.push_back(x);
  }
  return in;
}

int run_graph(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) return 2;
  int nv = 0, ne = 0;
  if (!read_all(f, &nv, 4) || !read_all(f, &ne, 4) || nv < 1 || ne < 0) return 2;
  std::vector<int> edges(2 * ne), inc_ptr(nv + 1), inc_edge(2 * ne), chi(ne), phys(nv);
  const bool ok = read_all(f, edges.data(), 8 * (size_t)ne) && read_all(f, inc_ptr.data(), 4 * (size_t)(nv + 1)) &&
                  read_all(f, inc_edge.data(), 8 * (size_t)ne) && read_all(f, chi.data(), 4 * (size_t)ne) &&
                  read_all(f, phys.data(), 4 * (size_t)nv);
  fclose(f);
  if (!ok) return 2;
  const Topo t{nv, ne, edges.data(), inc_ptr.data(), inc_edge.data()};

  // layout: contiguous, in order, sizes from the dimensions
  Layout lay;
  std::string err;
  const int rc = describe(t, chi.data(), phys.data(), kLim, lay, err);
  expect(rc == ITNBP_OK, "describe: " + err);
  if (rc != ITNBP_OK) return 1;
  long long mo = 0, po = 0;
  for (int m = 0; m < 2 * ne; ++m) {
    expect(lay.msg_off[m] == mo, "message offset " + std::to_string(m));
    mo += (long long)chi[m / 2] * chi[m / 2];
  }
  expect(lay.msg_total == mo, "message total");
  for (int v = 0; v < nv; ++v) {
    long long D = phys[v];
    for (int a = inc_ptr[v]; a < inc_ptr[v + 1]; ++a) D *= chi[inc_edge[a]];
    expect(lay.site_D[v] == D && lay.psi_off[v] == po, "site layout " + std::to_string(v));
    po += D;
  }
  expect(lay.psi_total == po, "site total");

  // parallel sequence: every edge, then every reversed edge
  std::vector<int> par;
  default_sequence(t, true, par);
  expect((int)par.size() == 2 * ne, "parallel sequence length");
  for (int e = 0; e < ne && (int)par.size() == 2 * ne; ++e)
    expect(par[e] == 2 * e && par[ne + e] == 2 * e + 1, "parallel sequence order");

  // forest cover: a permutation of the message ids; within the forest an id
  // belongs to, the messages it reads from the same forest come before it
  std::vector<int> fc;
  default_sequence(t, false, fc);
  std::vector<int> pos(2 * ne, -1);
  for (int i = 0; i < (int)fc.size(); ++i) {
    const int m = fc[i];
    expect(m >= 0 && m < 2 * ne && pos[m] < 0, "forest cover visits each message once");
    if (m >= 0 && m < 2 * ne) pos[m] = i;
  }
  expect((int)fc.size() == 2 * ne, "forest cover length");
  // forests are consecutive blocks; find each message's block by scanning for
  // the point where the edge set restarts (a reversed message is followed by
  // a forward one of a new forest only at block ends, so use the reverse pairing:
  // block [b, b + 2n) holds n forward ids then their n reverses in reverse order)
  {
    int b = 0;
    while (b < (int)fc.size()) {
      int n = 1;
      // the smallest n with fc[b + 2n - 1 - i] == fc[b + i] ^ 1 for all i < n and
      // the forward half closed under its own edges
      for (; b + 2 * n <= (int)fc.size(); ++n) {
        bool mirror = true;
        for (int i = 0; i < n && mirror; ++i) mirror = fc[b + 2 * n - 1 - i] == (fc[b + i] ^ 1);
        if (mirror) break;
      }
      expect(b + 2 * n <= (int)fc.size(), "forest cover block structure");
      if (b + 2 * n > (int)fc.size()) break;
      std::vector<char> in_block(2 * ne, 0);
      for (int i = b; i < b + 2 * n; ++i) in_block[fc[i]] = 1;
      for (int i = b; i < b + 2 * n; ++i)
        for (int x : inputs_of(t, fc[i]))
          if (in_block[x]) expect(pos[x] < i, "forest cover reads its own forest in order");
      b += 2 * n;
    }
  }

  // sequential runs: a partition of the sequence into maximal runs of
  // independent messages
  for (const std::vector<int>* seq : {&fc, &par}) {
    const auto runs = sequential_runs(t, *seq);
    int next = 0;
    for (size_t r = 0; r < runs.size(); ++r) {
      const int b = runs[r].first, e = runs[r].second;
      expect(b == next && e > b, "runs are consecutive and non-empty");
      next = e;
      for (int i = b; i < e; ++i)
        for (int j = b; j < e; ++j) {
          if (i == j) continue;
          bool reads_j = false;
          for (int x : inputs_of(t, (*seq)[i])) reads_j = reads_j || x == (*seq)[j];
          expect(!reads_j && (*seq)[i] != (*seq)[j], "a run holds no dependency");
        }
      if (r + 1 < runs.size()) {
        // maximal: the first message of the next run conflicts with this one
        const int m = (*seq)[e];
        bool conflict = false;
        for (int i = b; i < e; ++i) {
          const int x = (*seq)[i];
          if (x == m) conflict = true;
          for (int y : inputs_of(t, m)) conflict = conflict || y == x;
          for (int y : inputs_of(t, x)) conflict = conflict || y == m;
        }
        expect(conflict, "runs are maximal");
      }
    }
    expect(next == (int)seq->size(), "runs cover the sequence");
  }

  // check_sequence: from unset messages a parallel sweep reads unset messages
  // wherever a vertex has degree >= 2; every message set: fine; an id out of
  // range: ITNBP_ERR_ARG; need_old from an unset state: ITNBP_ERR_STATE
  std::vector<char> none(2 * ne, 0), all(2 * ne, 1);
  bool deg2 = false;
  for (int v = 0; v < nv; ++v) deg2 = deg2 || inc_ptr[v + 1] - inc_ptr[v] >= 2;
  expect(check_sequence(t, none, par, false, false, err) == (deg2 ? ITNBP_ERR_STATE : ITNBP_OK), "unset parallel");
  expect(check_sequence(t, all, par, false, true, err) == ITNBP_OK, "set parallel");
  expect(check_sequence(t, all, {2 * ne}, false, false, err) == ITNBP_ERR_ARG, "id out of range");
  expect(check_sequence(t, all, {-1}, true, false, err) == ITNBP_ERR_ARG, "negative id");
  if (ne > 0) expect(check_sequence(t, none, fc, true, true, err) == ITNBP_ERR_STATE, "need_old from unset");

  printf("{\"checks\": %ld, \"failures\": %d, \"forest\": [", checks, failures);
  for (size_t i = 0; i < fc.size(); ++i) printf("%s%d", i ? ", " : "", fc[i]);
  printf("]}\n");
  return failures ? 1 : 0;
}

// a description with one field broken; returns describe()'s code
struct Desc {
  int nv, ne;
  std::vector<int> edges, inc_ptr, inc_edge, chi, phys;
  int run(const Limits& lim = kLim) const {
    Layout lay;
    std::string err;
    const Topo t{nv, ne, edges.empty() ? nullptr : edges.data(), inc_ptr.empty() ? nullptr : inc_ptr.data(),
                 inc_edge.empty() ? nullptr : inc_edge.data()};
    return describe(t, chi.empty() ? nullptr : chi.data(), phys.empty() ? nullptr : phys.data(), lim, lay, err);
  }
};

// path 0 - 1 - 2, chi 2, d 2
Desc path3() { return Desc{3, 2, {0, 1, 1, 2}, {0, 1, 3, 4}, {0, 0, 1, 1}, {2, 2}, {2, 2, 2}}; }

int run_invalid() {
  std::vector<std::pair<std::string, int>> got;
  auto add = [&](const char* name, int code) { got.push_back({name, code}); };
  add("valid", path3().run());
  { Desc d = path3(); d.nv = 0; add("no_vertices", d.run()); }
  { Desc d = path3(); d.edges = {0, 1, 1, 1}; add("self_loop", d.run()); }
  { Desc d = path3(); d.edges = {0, 1, 1, 3}; add("vertex_out_of_range", d.run()); }
  { Desc d = path3(); d.inc_ptr = {0, 1, 3, 3}; add("inc_ptr_end", d.run()); }
  { Desc d = path3(); d.inc_ptr = {0, 3, 1, 4}; add("inc_ptr_decreasing", d.run()); }
  { Desc d = path3(); d.inc_edge = {0, 0, 2, 1}; add("incident_out_of_range", d.run()); }
  { Desc d = path3(); d.inc_edge = {1, 0, 1, 1}; add("incident_not_touching", d.run()); }
  { Desc d = path3(); d.inc_edge = {0, 0, 0, 1}; add("incident_twice", d.run()); }
  { Desc d = path3(); d.chi = {0, 2}; add("chi_zero", d.run()); }
  { Desc d = path3(); d.chi = {65, 2}; add("chi_65", d.run()); }
  { Desc d = path3(); d.phys = {2, 9, 2}; add("phys_9", d.run()); }
  { Desc d = path3(); d.chi = {64, 64}; add("site_too_large", d.run(Limits{6, 64, 8, 8191})); }
  { Desc d = path3(); add("degree_above_limit", d.run(Limits{1, 64, 8, 1LL << 30})); }
  { Desc d = path3(); d.chi.clear(); add("null_bond_dims", d.run()); }
  {
    std::vector<LargeBatch> b;
    std::string err;
    add("open_dim_65", large_batches({1000}, {65}, 1 << 20, 8, b, err));
    // batches: each within the budget or a single task, covering the tasks in order
    std::vector<long long> D;
    std::vector<int> P;
    unsigned s = 12345;
    for (int i = 0; i < 200; ++i) {
      s = s * 1103515245u + 12345u;
      D.push_back(64 + (s >> 8) % 100000);
      P.push_back(1 + (s >> 4) % 64);
    }
    const size_t budget = 4 << 20;
    const int rc = large_batches(D, P, budget, 8, b, err);
    size_t next = 0;
    bool ok = rc == ITNBP_OK;
    for (const auto& x : b) {
      ok = ok && x.begin == next && x.count > 0;
      long long mD = 0;
      int mP2 = 1, mR = 1;
      for (size_t i = x.begin; i < x.begin + x.count; ++i) {
        mD = std::max(mD, D[i]);
        mP2 = std::max(mP2, P[i] * P[i]);
        mR = std::max(mR, (int)(D[i] / P[i]));
      }
      ok = ok && x.maxD == mD && x.maxP2 == mP2 && x.maxR == mR;
      ok = ok && (x.count == 1 || x.count * (size_t)(2 * mD * 8) <= budget);
      next += x.count;
    }
    add("large_batches_invariants", ok && next == D.size() ? 0 : 1);
  }
  printf("{");
  for (size_t i = 0; i < got.size(); ++i) printf("%s\"%s\": %d", i ? ", " : "", got[i].first.c_str(), got[i].second);
  printf("}\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: graph_driver <graph.bin> | --invalid\n");
    return 2;
  }
  if (std::string(argv[1]) == "--invalid") return run_invalid();
  return run_graph(argv[1]);
}
