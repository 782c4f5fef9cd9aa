// Host-side topology logic of the BP library (declared in itnbp_graph.hpp).
#include "itnbp_graph.hpp"

#include <algorithm>
#include <deque>

#include "../../include/itnbp.h"

namespace itnbp_graph {

int describe(const Topo& t, const int* chi, const int* phys, const Limits& lim, Layout& out, std::string& err) {
  const int nv = t.nv, ne = t.ne;
  if (nv <= 0 || ne < 0 || (ne > 0 && (!t.edges || !chi)) || !t.inc_ptr || (ne > 0 && !t.inc_edge) || !phys) {
    err = "null or empty network description";
    return ITNBP_ERR_ARG;
  }
  if (t.inc_ptr[0] != 0 || t.inc_ptr[nv] != 2 * ne) {
    err = "inc_ptr must start at 0 and end at 2*ne";
    return ITNBP_ERR_ARG;
  }
  for (int v = 0; v < nv; ++v)
    if (t.inc_ptr[v + 1] < t.inc_ptr[v]) {
      err = "inc_ptr must be non-decreasing";
      return ITNBP_ERR_ARG;
    }
  for (int e = 0; e < ne; ++e) {
    const int a = t.edges[2 * e], b = t.edges[2 * e + 1];
    if (a < 0 || a >= nv || b < 0 || b >= nv || a == b) {
      err = "bad edge";
      return ITNBP_ERR_ARG;
    }
    if (chi[e] < 1 || chi[e] > lim.max_chi) {
      err = "bond dimension must be in 1.." + std::to_string(lim.max_chi);
      return ITNBP_ERR_UNSUPPORTED;
    }
  }
  std::vector<int> seen(ne, 0);
  for (int v = 0; v < nv; ++v) {
    const int k = t.inc_ptr[v + 1] - t.inc_ptr[v];
    if (k > lim.max_k) {
      err = "vertex degree > " + std::to_string(lim.max_k) + " is not supported";
      return ITNBP_ERR_UNSUPPORTED;
    }
    if (phys[v] < 1 || phys[v] > lim.max_phys) {
      err = "physical dimension must be in 1.." + std::to_string(lim.max_phys);
      return ITNBP_ERR_UNSUPPORTED;
    }
    for (int a = t.inc_ptr[v]; a < t.inc_ptr[v + 1]; ++a) {
      const int e = t.inc_edge[a];
      if (e < 0 || e >= ne) {
        err = "incident edge out of range";
        return ITNBP_ERR_ARG;
      }
      if (t.edges[2 * e] != v && t.edges[2 * e + 1] != v) {
        err = "incident edge does not touch its vertex";
        return ITNBP_ERR_ARG;
      }
      seen[e]++;
    }
  }
  for (int e = 0; e < ne; ++e)
    if (seen[e] != 2) {
      err = "every edge must be incident to both endpoints once";
      return ITNBP_ERR_ARG;
    }
  out = Layout();
  out.msg_off.resize(2 * (size_t)ne);
  for (int m = 0; m < 2 * ne; ++m) {
    out.msg_off[m] = out.msg_total;
    out.msg_total += (long long)chi[m / 2] * chi[m / 2];
    out.max_chi = std::max(out.max_chi, chi[m / 2]);
  }
  out.psi_off.resize(nv);
  out.site_D.resize(nv);
  for (int v = 0; v < nv; ++v) {
    const int k = t.inc_ptr[v + 1] - t.inc_ptr[v];
    out.max_k = std::max(out.max_k, k);
    long long D = phys[v];
    for (int a = t.inc_ptr[v]; a < t.inc_ptr[v + 1]; ++a) {
      D *= chi[t.inc_edge[a]];
      if (D > lim.max_site) {
        err = "site tensor larger than " + std::to_string(lim.max_site) + " elements";
        return ITNBP_ERR_UNSUPPORTED;
      }
    }
    out.site_D[v] = D;
    out.psi_off[v] = out.psi_total;
    out.psi_total += D;
  }
  return ITNBP_OK;
}

void default_sequence(const Topo& t, bool parallel, std::vector<int>& seq) {
  seq.clear();
  const int ne = t.ne, nv = t.nv;
  if (parallel) {
    for (int e = 0; e < ne; ++e) seq.push_back(2 * e);
    for (int e = 0; e < ne; ++e) seq.push_back(2 * e + 1);
    return;
  }
  std::vector<char> remaining(ne, 1);
  int left = ne;
  while (left > 0) {
    std::vector<std::vector<std::pair<int, int>>> adj(nv);  // (neighbour, edge)
    for (int e = 0; e < ne; ++e)
      if (remaining[e]) {
        const int a = t.edges[2 * e], b = t.edges[2 * e + 1];
        adj[a].push_back({b, e});
        adj[b].push_back({a, e});
      }
    for (auto& l : adj) std::sort(l.begin(), l.end());
    std::vector<char> seen(nv, 0);
    std::vector<std::vector<std::pair<int, int>>> children(nv);  // (child, edge)
    std::vector<int> roots;
    std::vector<char> in_forest(ne, 0);
    for (int r = 0; r < nv; ++r) {
      if (seen[r] || adj[r].empty()) continue;
      seen[r] = 1;
      roots.push_back(r);
      std::deque<int> q{r};
      while (!q.empty()) {
        const int x = q.front();
        q.pop_front();
        for (auto& ye : adj[x])
          if (!seen[ye.first]) {
            seen[ye.first] = 1;
            children[x].push_back(ye);
            in_forest[ye.second] = 1;
            q.push_back(ye.first);
          }
      }
    }
    std::vector<int> parent_edge(nv, -1);
    for (int r : roots) {
      std::vector<int> post;  // directed ids child -> parent
      std::vector<std::pair<int, size_t>> st{{r, 0}};
      while (!st.empty()) {
        auto& top = st.back();
        const int x = top.first;
        if (top.second < children[x].size()) {
          const auto ce = children[x][top.second++];
          parent_edge[ce.first] = ce.second;
          st.push_back({ce.first, 0});
        } else {
          st.pop_back();
          if (parent_edge[x] >= 0) {
            const int e = parent_edge[x];
            const int parent = t.edges[2 * e] == x ? t.edges[2 * e + 1] : t.edges[2 * e];
            post.push_back(msg_id_into(t, e, parent));
          }
        }
      }
      for (int m : post) seq.push_back(m);
      for (int i = (int)post.size() - 1; i >= 0; --i) seq.push_back(post[i] ^ 1);
    }
    for (int e = 0; e < ne; ++e)
      if (in_forest[e]) {
        remaining[e] = 0;
        --left;
      }
  }
}

namespace {

// message reader = (a -> b) reads m = (u -> v) iff v == a and u != b
bool reads(const Topo& t, int reader, int m) {
  return msg_dst(t, m) == msg_src(t, reader) && msg_src(t, m) != msg_dst(t, reader);
}

}  // namespace

std::vector<std::pair<int, int>> sequential_runs(const Topo& t, const std::vector<int>& seq) {
  std::vector<std::pair<int, int>> runs;
  int start = 0;
  for (int i = 0; i < (int)seq.size(); ++i) {
    bool conflict = false;
    for (int j = start; j < i && !conflict; ++j)
      conflict = seq[j] == seq[i] || reads(t, seq[i], seq[j]) || reads(t, seq[j], seq[i]);
    if (conflict) {
      runs.push_back({start, i});
      start = i;
    }
  }
  if (start < (int)seq.size()) runs.push_back({start, (int)seq.size()});
  return runs;
}

int check_sequence(const Topo& t, const std::vector<char>& is_set, const std::vector<int>& seq, bool sequential,
                   bool need_old, std::string& err) {
  std::vector<char> set(is_set);
  set.resize(2 * (size_t)t.ne, 0);
  for (int m : seq) {
    if (m < 0 || m >= 2 * t.ne) {
      err = "message id out of range in edge sequence";
      return ITNBP_ERR_ARG;
    }
    const int u = msg_src(t, m), v = msg_dst(t, m);
    for (int a = t.inc_ptr[u]; a < t.inc_ptr[u + 1]; ++a) {
      const int in = msg_id_into(t, t.inc_edge[a], u);
      if (msg_src(t, in) == v) continue;
      if (!set[in]) {
        err = "message " + std::to_string(in) + " is read before it is set (messages have no default)";
        return ITNBP_ERR_STATE;
      }
    }
    if (need_old && !set[m]) {
      err = "message " + std::to_string(m) + " has no previous value to compare with (tol needs one)";
      return ITNBP_ERR_STATE;
    }
    if (sequential) set[m] = 1;
  }
  return ITNBP_OK;
}

int large_batches(const std::vector<long long>& D, const std::vector<int>& P, size_t budget, int elem_bytes,
                  std::vector<LargeBatch>& out, std::string& err) {
  out.clear();
  const size_t n = D.size();
  if (P.size() != n) {
    err = "large_batches: size mismatch";
    return ITNBP_ERR_ARG;
  }
  size_t i0 = 0;
  while (i0 < n) {
    LargeBatch b;
    b.begin = i0;
    size_t i1 = i0;
    for (; i1 < n; ++i1) {
      const long long nD = std::max(b.maxD, D[i1]);
      const size_t need = (i1 - i0 + 1) * (size_t)(2 * nD * elem_bytes);
      if (i1 > i0 && need > budget) break;
      b.maxD = nD;
    }
    for (size_t i = i0; i < i1; ++i) {
      if (P[i] < 1 || P[i] > 64) {
        err = "open dimension > 64 is not supported";
        return ITNBP_ERR_UNSUPPORTED;
      }
      b.maxP2 = std::max(b.maxP2, P[i] * P[i]);
      b.maxR = std::max(b.maxR, (int)(D[i] / P[i]));
    }
    b.count = i1 - i0;
    out.push_back(b);
    i0 = i1;
  }
  return ITNBP_OK;
}

}  // namespace itnbp_graph
