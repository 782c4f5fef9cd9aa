// Host-side topology logic of the BP library, free of HIP so that it builds
// with a plain C++ compiler (and, in the tests, under the address and
// undefined-behaviour sanitizers):
//   - validation of a network description and the storage layout of the site
//     tensors and messages;
//   - the default edge sequences (edge_sequences.jl:9-51);
//   - the split of an edge sequence into runs of mutually independent
//     messages, and the check that a sequence reads no unset message;
//   - the batching of large contraction tasks under a scratch budget.
#pragma once

#include <stddef.h>

#include <string>
#include <utility>
#include <vector>

namespace itnbp_graph {

// A network's graph as the C ABI describes it: edge e joins edges[2e] and
// edges[2e+1]; the incident edges of vertex v are inc_edge[inc_ptr[v] ..
// inc_ptr[v+1]), in the order of the site tensor's bond axes.  Message id 2e
// runs edges[2e] -> edges[2e+1], id 2e+1 the other way.
struct Topo {
  int nv = 0, ne = 0;
  const int* edges = nullptr;
  const int* inc_ptr = nullptr;
  const int* inc_edge = nullptr;
};

inline int msg_id_into(const Topo& t, int e, int to) { return t.edges[2 * e + 1] == to ? 2 * e : 2 * e + 1; }
inline int msg_src(const Topo& t, int m) { return t.edges[2 * (m / 2) + (m & 1)]; }
inline int msg_dst(const Topo& t, int m) { return t.edges[2 * (m / 2) + 1 - (m & 1)]; }

struct Limits {
  int max_k;              // vertex degree
  int max_chi;            // bond dimension
  int max_phys;           // physical dimension
  long long max_site;     // elements of one site tensor
};

// Storage layout: messages chi_e^2 doubles each in id order, site tensors
// column-major [s, bonds in incident order] one after another.
struct Layout {
  std::vector<long long> msg_off;  // [2 ne]
  long long msg_total = 0;
  std::vector<long long> psi_off;  // [nv]
  std::vector<long long> site_D;   // [nv]
  long long psi_total = 0;
  int max_chi = 0, max_k = 0;
};

// Checks the description and computes its layout.  Returns ITNBP_OK, or an
// ITNBP_ERR_* code with the reason in err.
int describe(const Topo& t, const int* chi, const int* phys, const Limits& lim, Layout& out, std::string& err);

// The default sequence of update(bpc): for the parallel schedule every edge
// then every reversed edge (edge_sequences.jl:49-51); for the sequential one
// the forest cover: spanning forests of the edges left over, each tree rooted
// at its lowest vertex and traversed breadth first, its edges in post order
// toward the root followed by their reverses (edge_sequences.jl:32-47).
void default_sequence(const Topo& t, bool parallel, std::vector<int>& seq);

// Consecutive runs [begin, end) of seq in which no message reads another one
// of the run (or repeats it): each run can be updated in one launch and still
// give the result of the one-by-one sequential update.
std::vector<std::pair<int, int>> sequential_runs(const Topo& t, const std::vector<int>& seq);

// The errors update(bpc) raises for a sequence: an id out of range
// (ITNBP_ERR_ARG), an incoming message read before it was set, or, with
// need_old, a message without a previous value (ITNBP_ERR_STATE).  is_set has
// one entry per message id; sequential: a message counts as set once the
// sequence has updated it, else only after the whole sweep.
int check_sequence(const Topo& t, const std::vector<char>& is_set, const std::vector<int>& seq, bool sequential,
                   bool need_old, std::string& err);

// Consecutive batches of large contraction tasks (site size D[i], open
// dimension P[i]) whose scratch of 2 * maxD * elem_bytes per task fits in
// budget (a batch holds at least one task).  ITNBP_ERR_UNSUPPORTED for an open
// dimension above 64.
struct LargeBatch {
  size_t begin = 0, count = 0;
  long long maxD = 0;
  int maxP2 = 1;  // largest P^2
  int maxR = 1;   // largest D / P
};
int large_batches(const std::vector<long long>& D, const std::vector<int>& P, size_t budget, int elem_bytes,
                  std::vector<LargeBatch>& out, std::string& err);

}  // namespace itnbp_graph
