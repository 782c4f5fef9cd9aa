// C++ caller of the C ABI (include/itnbp.h): builds a network from a JSON-free
// binary input, runs BP and the observables, and prints the results as one
// JSON object for the test that compares them with the oracle.
//
//   abi_driver <input.bin>
//
// Input (little-endian): int32 nv, ne; int32 edges[2 ne]; int32 inc_ptr[nv + 1];
// int32 inc_edge[2 ne]; int32 chi[ne]; int32 phys[nv]; float64 psi[...]
// (column-major [s, bonds in incident order], vertex after vertex).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "itnbp.h"

namespace {

bool read_all(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

int fail(itnbp_ctx* ctx, const char* what, int rc) {
  fprintf(stderr, "%s failed (%d): %s\n", what, rc, ctx ? itnbp_last_error(ctx) : "");
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: abi_driver <input.bin>\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int nv = 0, ne = 0;
  if (!read_all(f, &nv, 4) || !read_all(f, &ne, 4)) return 2;
  std::vector<int> edges(2 * ne), inc_ptr(nv + 1), inc_edge(2 * ne), chi(ne), phys(nv);
  bool ok = read_all(f, edges.data(), 8 * (size_t)ne) && read_all(f, inc_ptr.data(), 4 * (size_t)(nv + 1)) &&
            read_all(f, inc_edge.data(), 8 * (size_t)ne) && read_all(f, chi.data(), 4 * (size_t)ne) &&
            read_all(f, phys.data(), 4 * (size_t)nv);
  if (!ok) return 2;
  size_t psi_n = 0;
  for (int v = 0; v < nv; ++v) {
    size_t D = phys[v];
    for (int a = inc_ptr[v]; a < inc_ptr[v + 1]; ++a) D *= chi[inc_edge[a]];
    psi_n += D;
  }
  std::vector<double> psi(psi_n);
  if (!read_all(f, psi.data(), 8 * psi_n)) return 2;
  fclose(f);

  itnbp_ctx* ctx = nullptr;
  int rc = itnbp_create(0, &ctx);
  if (rc != ITNBP_OK) return fail(nullptr, "itnbp_create", rc);
  itnbp_net* net = nullptr;
  rc = itnbp_net_create(ctx, nv, ne, edges.data(), inc_ptr.data(), inc_edge.data(), chi.data(), phys.data(),
                        ITNBP_F64, &net);
  if (rc != ITNBP_OK) return fail(ctx, "itnbp_net_create", rc);
  if ((size_t)itnbp_net_psi_numel(net) != psi_n) return fail(ctx, "itnbp_net_psi_numel", -1);
  if ((rc = itnbp_set_sites(net, psi.data())) != ITNBP_OK) return fail(ctx, "itnbp_set_sites", rc);
  if ((rc = itnbp_init_identity_messages(net)) != ITNBP_OK) return fail(ctx, "itnbp_init_identity_messages", rc);

  // the errors the reference raises: no maxiter (abstractbeliefpropagationcache.jl:315-317)
  itnbp_update_opts bad = {ITNBP_SCHED_SEQUENTIAL, 0, -1, -1, ITNBP_PREC_F64, 0, 1, 0, nullptr, 0};
  itnbp_update_result r{};
  const int rc_bad = itnbp_update(net, &bad, &r);

  // sequential schedule (forest cover), 30 sweeps, tol 1e-12
  itnbp_update_opts o = {ITNBP_SCHED_SEQUENTIAL, 30, 1e-12, -1, ITNBP_PREC_F64, 0, 1, 0, nullptr, 0};
  if ((rc = itnbp_update(net, &o, &r)) != ITNBP_OK) return fail(ctx, "itnbp_update", rc);
  std::vector<double> msgs(itnbp_net_msg_numel(net));
  if ((rc = itnbp_get_messages(net, msgs.data())) != ITNBP_OK) return fail(ctx, "itnbp_get_messages", rc);
  std::vector<int> verts(nv);
  std::vector<double> sz_ops, sz(nv);
  for (int v = 0; v < nv; ++v) {
    verts[v] = v;
    for (int j = 0; j < phys[v]; ++j)
      for (int i = 0; i < phys[v]; ++i) sz_ops.push_back(i == j ? (i == 0 ? 0.5 : (i == 1 ? -0.5 : 0.0)) : 0.0);
  }
  if ((rc = itnbp_expect1(net, nv, verts.data(), sz_ops.data(), sz.data())) != ITNBP_OK)
    return fail(ctx, "itnbp_expect1", rc);
  double re = 0, im = 0;
  int ninf = 0;
  if ((rc = itnbp_logscalar(net, &re, &im, &ninf)) != ITNBP_OK) return fail(ctx, "itnbp_logscalar", rc);
  std::vector<int> seq(2 * ne);
  if ((rc = itnbp_forest_cover_sequence(net, seq.data())) != ITNBP_OK) return fail(ctx, "itnbp_forest_cover_sequence", rc);

  printf("{\"iterations\": %d, \"mean_diff\": %.17g, \"no_maxiter_rc\": %d, \"logscalar\": [%.17g, %.17g, %d],\n",
         r.iterations, r.mean_diff, rc_bad, re, im, ninf);
  printf(" \"sz\": [");
  for (int v = 0; v < nv; ++v) printf("%s%.17g", v ? ", " : "", sz[v]);
  printf("],\n \"sequence\": [");
  for (int i = 0; i < 2 * ne; ++i) printf("%s%d", i ? ", " : "", seq[i]);
  printf("],\n \"messages\": [");
  for (size_t i = 0; i < msgs.size(); ++i) printf("%s%.17g", i ? ", " : "", msgs[i]);
  printf("]}\n");
  itnbp_net_destroy(net);
  itnbp_destroy(ctx);
  return 0;
}
