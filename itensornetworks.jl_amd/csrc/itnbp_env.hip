// Environments, expectation values, region scalars, logscalar and
// normalisation on the device.
//
//   environment(bpc, verts)      src/caches/beliefpropagationcache.jl:100-105
//   expect(psiIpsi, op)          src/expect.jl:5-19 (numerator / denominator)
//   two-site RDM pattern         test/test_belief_propagation.jl:64-91
//   region_scalar                src/caches/beliefpropagationcache.jl:107-119
//   logscalar                    src/caches/abstractbeliefpropagationcache.jl:397-408
//   rescale_messages             src/caches/beliefpropagationcache.jl:121-139
//   rescale_partitions           src/caches/abstractbeliefpropagationcache.jl:349-379
//   normalize(alg="bp")          src/normalize.jl:125-142
#include <cmath>
#include <complex>
#include <vector>

#include "itnbp_ops.hpp"

namespace {

int check_all_incoming_set(itnbp_net* net, int v, int skip_e) {
  for (int a = net->inc_ptr[v]; a < net->inc_ptr[v + 1]; ++a) {
    const int e = net->inc_edge[a];
    if (e == skip_e) continue;
    if (!net->msg_set[msg_id_into(net, e, v)])
      return itnbp_fail(net->ctx, ITNBP_ERR_STATE, "incoming message of vertex " + std::to_string(v) + " is not set");
  }
  return ITNBP_OK;
}

// rho_v for the listed vertices into aux (d_v^2 each), offsets returned.  Sites
// of the chi = 32 eigenbasis pipeline take the rotated-Gram route, the others
// the generic open contraction.
int rdm1_device(itnbp_net* net, int n, const int* verts, std::vector<long long>& offs, long long& tot) {
  tot = 0;
  offs.clear();
  for (int i = 0; i < n; ++i) {
    const int v = verts[i];
    if (v < 0 || v >= net->nv) return itnbp_fail(net->ctx, ITNBP_ERR_ARG, "vertex out of range");
    ITNBP_RETCHK(check_all_incoming_set(net, v, -1));
    offs.push_back(tot);
    tot += (long long)net->phys[v] * net->phys[v];
  }
  ITNBP_RETCHK(ensure_buf(net->ctx, net->aux, std::max<long long>(1, tot) * 8));
  std::vector<char> done;
  ITNBP_RETCHK(fast_site_rdm(net, n, verts, offs.data(), (double*)net->aux.p, done));
  std::vector<OcTask> tasks;
  for (int i = 0; i < n; ++i) {
    if (done[i]) continue;
    OcTask t;
    t.v = verts[i];
    t.j = 0;
    t.s_open = 1;
    t.absorb_mask = all_bonds_mask(net, t.v);
    t.out_off = offs[i];
    t.old_off = -1;
    t.normalize = 0;
    t.slot = (int)tasks.size();
    tasks.push_back(t);
  }
  if (tasks.empty()) return ITNBP_OK;
  ITNBP_RETCHK(ensure_buf(net->ctx, net->stats, sizeof(MsgStat) * std::max<size_t>(1, tasks.size())));
  return run_tasks(net, tasks, 1, net->d_psi, net->dtype, net->d_msg[net->cur], net->aux.p,
                   (MsgStat*)net->stats.p, net->tasks);
}

}  // namespace

extern "C" {

int itnbp_rdm1(itnbp_net* net, int n, const int* verts, double* out) {
  if (!net || n < 0 || (n > 0 && (!verts || !out))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  std::vector<long long> offs;
  long long tot = 0;
  ITNBP_RETCHK(rdm1_device(net, n, verts, offs, tot));
  ITNBP_HIPCHK(net->ctx, hipMemcpyAsync(out, net->aux.p, tot * 8, hipMemcpyDeviceToHost, net->ctx->stream));
  ITNBP_HIPCHK(net->ctx, hipStreamSynchronize(net->ctx->stream));
  return ITNBP_OK;
}

int itnbp_expect1(itnbp_net* net, int n, const int* verts, const double* ops, double* out) {
  if (!net || n < 0 || (n > 0 && (!verts || !ops || !out))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  std::vector<long long> offs;
  long long tot = 0;
  ITNBP_RETCHK(rdm1_device(net, n, verts, offs, tot));
  std::vector<double> rho(tot);
  ITNBP_HIPCHK(net->ctx, hipMemcpyAsync(rho.data(), net->aux.p, tot * 8, hipMemcpyDeviceToHost, net->ctx->stream));
  ITNBP_HIPCHK(net->ctx, hipStreamSynchronize(net->ctx->stream));
  long long oo = 0;
  for (int i = 0; i < n; ++i) {
    const int d = net->phys[verts[i]];
    const double* r = rho.data() + offs[i];
    const double* O = ops + oo;
    double num = 0, den = 0;
    // sum_{s,s'} O[s', s] rho[s, s'] / sum_s rho[s, s]
    for (int s = 0; s < d; ++s) {
      den += r[s + d * s];
      for (int sp = 0; sp < d; ++sp) num += O[sp + d * s] * r[s + d * sp];
    }
    out[i] = num / den;
    oo += (long long)d * d;
  }
  return ITNBP_OK;
}

int itnbp_rdm2(itnbp_net* net, int n, const int* edge_ids, double* out) {
  if (!net || n < 0 || (n > 0 && (!edge_ids || !out))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  itnbp_ctx* ctx = net->ctx;
  std::vector<OcTask> tasks;
  std::vector<long long> au, av, oo;
  std::vector<int> dus, dvs, chis;
  long long tot = 0, otot = 0;
  for (int i = 0; i < n; ++i) {
    const int e = edge_ids[i];
    if (e < 0 || e >= net->ne) return itnbp_fail(ctx, ITNBP_ERR_ARG, "edge id out of range");
    const int u = net->edges[2 * e], v = net->edges[2 * e + 1];
    ITNBP_RETCHK(check_all_incoming_set(net, u, e));
    ITNBP_RETCHK(check_all_incoming_set(net, v, e));
    const int chi = net->chi[e];
    for (int side = 0; side < 2; ++side) {
      const int x = side == 0 ? u : v;
      OcTask t;
      t.v = x;
      t.j = axis_of(net, x, e);
      t.s_open = 1;
      t.absorb_mask = all_bonds_mask(net, x) & ~(1 << t.j);
      t.out_off = tot;
      t.old_off = -1;
      t.normalize = 0;
      t.slot = 2 * i + side;
      (side == 0 ? au : av).push_back(tot);
      const long long L = (long long)net->phys[x] * chi;
      tot += L * L;
      tasks.push_back(t);
    }
    dus.push_back(net->phys[u]);
    dvs.push_back(net->phys[v]);
    chis.push_back(chi);
    oo.push_back(otot);
    otot += (long long)net->phys[u] * net->phys[v] * net->phys[u] * net->phys[v];
  }
  ITNBP_RETCHK(ensure_buf(ctx, net->aux, (tot + otot) * 8));
  // half-pairs of chi = 32 sites from the eigenbasis rotations, the rest as
  // generic open contractions
  {
    std::vector<int> hs(2 * n), ha(2 * n);
    std::vector<long long> ho(2 * n);
    for (size_t k = 0; k < tasks.size(); ++k) {
      hs[k] = tasks[k].v;
      ha[k] = tasks[k].j;
      ho[k] = tasks[k].out_off;
    }
    std::vector<char> done;
    ITNBP_RETCHK(fast_pair_halves(net, 2 * n, hs.data(), ha.data(), ho.data(), (double*)net->aux.p, done));
    std::vector<OcTask> rest;
    for (size_t k = 0; k < tasks.size(); ++k)
      if (!done[k]) {
        rest.push_back(tasks[k]);
        rest.back().slot = (int)rest.size() - 1;
      }
    tasks.swap(rest);
  }
  if (!tasks.empty()) {
    ITNBP_RETCHK(ensure_buf(ctx, net->stats, sizeof(MsgStat) * tasks.size()));
    ITNBP_RETCHK(run_tasks(net, tasks, 1, net->d_psi, net->dtype, net->d_msg[net->cur], net->aux.p,
                           (MsgStat*)net->stats.p, net->tasks));
  }
  // small index tables
  DevBuf tab;
  const size_t tb = n * (2 * sizeof(long long) + 3 * sizeof(int));
  ITNBP_RETCHK(ensure_buf(ctx, tab, tb));
  char* base = (char*)tab.p;
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base, au.data(), n * 8, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + n * 8, av.data(), n * 8, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + n * 16, dus.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + n * 20, dvs.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + n * 24, chis.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  // outputs assume uniform (d_u d_v)^2 stride only through oo; the kernel uses p*dd*dd
  bool uniform = true;
  for (int i = 1; i < n; ++i) uniform = uniform && dus[i] == dus[0] && dvs[i] == dvs[0];
  double* d_out = (double*)net->aux.p + tot;
  if (uniform) {
    ITNBP_HIPCHK(ctx, launch_rdm2_combine((const double*)net->aux.p, (const long long*)base,
                                          (const long long*)(base + n * 8), (const int*)(base + n * 16),
                                          (const int*)(base + n * 20), (const int*)(base + n * 24), n,
                                          d_out, ctx->stream));
  } else {
    for (int i = 0; i < n; ++i)
      ITNBP_HIPCHK(ctx, launch_rdm2_combine((const double*)net->aux.p, (const long long*)base + i,
                                            (const long long*)(base + n * 8) + i,
                                            (const int*)(base + n * 16) + i, (const int*)(base + n * 20) + i,
                                            (const int*)(base + n * 24) + i, 1, d_out + oo[i], ctx->stream));
  }
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(out, d_out, otot * 8, hipMemcpyDeviceToHost, ctx->stream));
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_buf(tab);
  return ITNBP_OK;
}

int itnbp_expect2(itnbp_net* net, int n, const int* edge_ids, const double* ops, double* out) {
  if (!net || n < 0 || (n > 0 && (!edge_ids || !ops || !out))) return ITNBP_ERR_ARG;
  if (n == 0) return ITNBP_OK;
  long long tot = 0;
  for (int i = 0; i < n; ++i) {
    if (edge_ids[i] < 0 || edge_ids[i] >= net->ne) return itnbp_fail(net->ctx, ITNBP_ERR_ARG, "edge id out of range");
    const long long dd = (long long)net->phys[net->edges[2 * edge_ids[i]]] * net->phys[net->edges[2 * edge_ids[i] + 1]];
    tot += dd * dd;
  }
  std::vector<double> rho(tot);
  ITNBP_RETCHK(itnbp_rdm2(net, n, edge_ids, rho.data()));
  long long oo = 0;
  for (int i = 0; i < n; ++i) {
    const int e = edge_ids[i];
    const int dd = net->phys[net->edges[2 * e]] * net->phys[net->edges[2 * e + 1]];
    const double* r = rho.data() + oo;
    const double* O = ops + oo;
    double num = 0, den = 0;
    for (int s = 0; s < dd; ++s) {
      den += r[s + dd * s];
      for (int sp = 0; sp < dd; ++sp) num += O[sp + dd * s] * r[s + dd * sp];
    }
    out[i] = num / den;
    oo += (long long)dd * dd;
  }
  return ITNBP_OK;
}

int itnbp_region_scalars(itnbp_net* net, double* vs, double* es) {
  if (!net) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  if (vs) {
    // vertex scalar = tr rho_v: sites of the chi = 32 eigenbasis pipeline from
    // their density matrices (aux[nv + 4 v]), the others as closed contractions
    std::vector<int> all(net->nv);
    std::vector<long long> roff(net->nv);
    for (int v = 0; v < net->nv; ++v) {
      ITNBP_RETCHK(check_all_incoming_set(net, v, -1));
      all[v] = v;
      roff[v] = net->nv + 4LL * v;
    }
    ITNBP_RETCHK(ensure_buf(ctx, net->aux, std::max(1, 5 * net->nv) * 8));
    std::vector<char> done;
    ITNBP_RETCHK(fast_site_rdm(net, net->nv, all.data(), roff.data(), (double*)net->aux.p, done));
    std::vector<OcTask> tasks;
    for (int v = 0; v < net->nv; ++v) {
      if (done[v]) continue;
      OcTask t;
      t.v = v;
      t.j = 0;
      t.s_open = 0;
      t.absorb_mask = all_bonds_mask(net, v);
      t.out_off = v;
      t.old_off = -1;
      t.normalize = 0;
      t.slot = (int)tasks.size();
      tasks.push_back(t);
    }
    if (!tasks.empty()) {
      ITNBP_RETCHK(ensure_buf(ctx, net->stats, sizeof(MsgStat) * tasks.size()));
      ITNBP_RETCHK(run_tasks(net, tasks, 1, net->d_psi, net->dtype, net->d_msg[net->cur], net->aux.p,
                             (MsgStat*)net->stats.p, net->tasks));
    }
    std::vector<double> h(5 * (size_t)net->nv);
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(h.data(), net->aux.p, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    for (int v = 0; v < net->nv; ++v) vs[v] = done[v] ? h[net->nv + 4LL * v] + h[net->nv + 4LL * v + 3] : h[v];
  }
  if (es && net->ne > 0) {
    for (int m = 0; m < 2 * net->ne; ++m)
      if (!net->msg_set[m]) return itnbp_fail(ctx, ITNBP_ERR_STATE, "message " + std::to_string(m) + " is not set");
    DevBuf tab;
    const size_t chi_at = 16 * (size_t)net->ne;
    const size_t out_at = chi_at + ((4 * (size_t)net->ne + 15) / 16) * 16;
    const size_t tb = out_at + 8 * (size_t)net->ne;
    ITNBP_RETCHK(ensure_buf(ctx, tab, tb));
    char* base = (char*)tab.p;
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(base, net->msg_off.data(), 2 * net->ne * 8, hipMemcpyHostToDevice, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + 16 * net->ne, net->chi.data(), net->ne * 4, hipMemcpyHostToDevice, ctx->stream));
    double* d_out = (double*)(base + out_at);
    ITNBP_HIPCHK(ctx, launch_edge_scalars(net->d_msg[net->cur], (const long long*)base,
                                          (const int*)(base + 16 * net->ne), net->ne, d_out, ctx->stream));
    ITNBP_HIPCHK(ctx, hipMemcpyAsync(es, d_out, net->ne * 8, hipMemcpyDeviceToHost, ctx->stream));
    ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    free_buf(tab);
  }
  return ITNBP_OK;
}

int itnbp_logscalar(itnbp_net* net, double* re, double* im, int* neg_inf) {
  if (!net || !re || !im || !neg_inf) return ITNBP_ERR_ARG;
  std::vector<double> vs(net->nv), es(std::max(1, net->ne));
  ITNBP_RETCHK(itnbp_region_scalars(net, vs.data(), es.data()));
  *neg_inf = 0;
  for (int e = 0; e < net->ne; ++e)
    if (es[e] == 0.0) {
      *neg_inf = 1;
      *re = -INFINITY;
      *im = 0;
      return ITNBP_OK;
    }
  // complex promotion when any term is negative (abstractbeliefpropagationcache.jl:399-404)
  std::complex<double> acc(0.0, 0.0);
  bool vneg = false, eneg = false;
  for (double x : vs) vneg = vneg || x < 0;
  for (int e = 0; e < net->ne; ++e) eneg = eneg || es[e] < 0;
  for (double x : vs) acc += vneg ? std::log(std::complex<double>(x, 0.0)) : std::complex<double>(std::log(x), 0.0);
  for (int e = 0; e < net->ne; ++e)
    acc -= eneg ? std::log(std::complex<double>(es[e], 0.0)) : std::complex<double>(std::log(es[e]), 0.0);
  *re = acc.real();
  *im = acc.imag();
  return ITNBP_OK;
}

int itnbp_normalize(itnbp_net* net) {
  if (!net) return ITNBP_ERR_ARG;
  itnbp_ctx* ctx = net->ctx;
  for (int m = 0; m < 2 * net->ne; ++m)
    if (!net->msg_set[m]) return itnbp_fail(ctx, ITNBP_ERR_STATE, "normalize needs every message set");
  // rescale_messages: normalise both messages of an edge, fix the sign of the
  // edge scalar, scale both by Z_e^{-1/2}
  std::vector<double> h(net->msg_total);
  ITNBP_RETCHK(itnbp_get_messages(net, h.data()));
  for (int e = 0; e < net->ne; ++e) {
    const long long n2 = (long long)net->chi[e] * net->chi[e];
    double* a = h.data() + net->msg_off[2 * e];
    double* b = h.data() + net->msg_off[2 * e + 1];
    double na = 0, nb = 0;
    for (long long i = 0; i < n2; ++i) {
      na += a[i] * a[i];
      nb += b[i] * b[i];
    }
    na = std::sqrt(na);
    nb = std::sqrt(nb);
    double z = 0;
    for (long long i = 0; i < n2; ++i) z += (a[i] / na) * (b[i] / nb);
    const double sg = z < 0 ? -1.0 : 1.0;
    z *= sg;
    const double sf = 1.0 / std::sqrt(z);
    for (long long i = 0; i < n2; ++i) {
      a[i] = sf * sg * a[i] / na;
      b[i] = sf * b[i] / nb;
    }
  }
  ITNBP_RETCHK(itnbp_set_messages(net, h.data()));
  // rescale_partitions: psi_v / ||psi_v||, then * Z_v^{-1/2}; equal to
  // psi_v * Z_v(psi_v)^{-1/2} since Z_v is quadratic in psi_v
  std::vector<double> vs(net->nv);
  ITNBP_RETCHK(itnbp_region_scalars(net, vs.data(), nullptr));
  std::vector<double> sc(net->nv);
  for (int v = 0; v < net->nv; ++v) sc[v] = vs[v] > 0 ? 1.0 / std::sqrt(vs[v]) : 1.0;
  std::vector<int> sizes(net->nv);
  for (int v = 0; v < net->nv; ++v) sizes[v] = net->vdesc[v].D;
  DevBuf tab;
  const size_t tb = (size_t)net->nv * (8 + 8 + 4);
  ITNBP_RETCHK(ensure_buf(ctx, tab, tb));
  char* base = (char*)tab.p;
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base, net->psi_off.data(), net->nv * 8, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + 8 * net->nv, sc.data(), net->nv * 8, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, hipMemcpyAsync(base + 16 * net->nv, sizes.data(), net->nv * 4, hipMemcpyHostToDevice, ctx->stream));
  ITNBP_HIPCHK(ctx, launch_scale_psi(net->dtype, net->d_psi, (const long long*)base, (const int*)(base + 16 * net->nv),
                                     (const double*)(base + 8 * net->nv), net->nv, ctx->stream));
  ITNBP_HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  free_buf(tab);
  invalidate_fast_state(net);
  return ITNBP_OK;
}

}  // extern "C"
