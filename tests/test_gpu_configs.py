"""GPU parity at the sizes of the benchmark configurations, against the CPU oracle.

* 16x16 open PEPS, chi = 8, FP64 N(0,1): Jacobi BP to mean message_diff 1e-10,
  then one BP-gauged round of exp(-0.05 H_b) on all 480 bonds, truncated to
  chi = 8 (the bulk bond sits exactly at the apply kernel's size limit,
  external size x d chi = 512 x 16).
* 256 heterogeneous FP32 instances (grids L in 6..12 and heavy-hex lattices,
  chi in {4, 6, 8, 12, 16}) in one batched update: the smallest instance of
  every (lattice kind, chi) class against the oracle after 10 sweeps, and
  every instance after 50 sweeps against the oracle's update map (fixed-point
  residual).
* 32x32 open PEPS, chi = 32, BF16 storage: one Jacobi sweep from identity
  messages in FP64, three-term BF16 and one-term 16-bit arithmetic, sampled
  messages of bulk, edge and corner sites against the oracle.
"""
import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, oracle_psi, to_index
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


def test_bp_16x16_chi8_to_1e10_and_gate_round_match_oracle():
    g = itnbp.named_grid((16, 16))
    psi = itnbp.random_tensornetwork(np.random.default_rng(2), np.float64, g, link_space=8)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    bpc = itnbp.initialize_cache(psi)
    itnbp.update_(bpc, maxiter=100, tol=1e-10, edge_sequence="parallel")
    it = bpc.last_update["iterations"]
    assert bpc.last_update["converged"]
    print(f"16x16 chi=8: {it} Jacobi sweeps to mean message_diff 1e-10, path {bpc.last_update['path']}", flush=True)
    ref, it_ref, md_ref, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=100, tol=1e-10,
                                         schedule="parallel")
    assert it == it_ref
    got = to_index(g, bpc.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    print(f"messages vs oracle: max abs err {err:.2e}", flush=True)
    assert err < 1e-10
    # message_diff is 1 - cos^2: near 1e-10 it carries an absolute rounding of
    # order 1e-16 per message (measured 1.6e-6 relative at 9.9e-11)
    assert bpc.last_update["mean_diff"] == pytest.approx(md_ref, rel=0, abs=2e-15)

    # one BP-gauged gate per bond, colour class by colour class, envs from this cache
    deg = {v: g.degree(v) for v in g.vertices}
    odeg = [len(og.incident[i]) for i in range(og.nv)]
    errs = itnbp.simple_update_round_(
        bpc, lambda u, v: O.expm_sym(O.ising_bond_hamiltonian(2.5, deg[u], deg[v]), -0.05), maxdim=8,
        normalize=True)
    assert len(errs) == 480
    ref_psi, ref_errs, _ = O.simple_update_round(
        og, opsi, got, lambda a, b: O.expm_sym(O.ising_bond_hamiltonian(2.5, odeg[a], odeg[b]), -0.05),
        maxdim=8, normalize=True)
    new = bpc.tensornetwork()
    terr = max(np.abs(new[v] - ref_psi[i]).max() for i, v in enumerate(g.vertices))
    print(f"gate round: max abs tensor err {terr:.2e}", flush=True)
    assert terr < 1e-10
    for (a, b), e in errs.items():
        r = ref_errs[(g.index[a], g.index[b])]
        assert e == pytest.approx(r, rel=1e-8, abs=1e-15)
    assert max(errs.values()) > 0


@pytest.fixture(scope="module")
def instance_set():
    return itnbp.random_instance_set(np.random.default_rng(3), 256)


def test_batched_256_instances_sampled_against_oracle(instance_set):
    inst = instance_set
    caches = [itnbp.initialize_cache(psi) for _, _, psi in inst]
    res = itnbp.update_batched_(caches, maxiter=10, precision="f32")
    assert all(r["iterations"] == 10 for r in res)
    # the smallest instance of every (lattice kind, chi) class
    pick = {}
    for i, (kind, chi, psi) in enumerate(inst):
        key = (kind, chi)
        if key not in pick or psi.graph.nv < inst[pick[key]][2].graph.nv:
            pick[key] = i
    assert len(pick) == 10, sorted(pick)
    worst = 0.0
    for (kind, chi), i in sorted(pick.items()):
        psi = inst[i][2]
        og, opsi = oracle_graph(psi.graph), oracle_psi(psi)
        ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=10, schedule="parallel")
        got = to_index(psi.graph, caches[i].messages())
        err = max(np.abs(got[e] - ref[e]).max() for e in ref)
        print(f"instance {i} ({kind}, chi={chi}, {psi.graph.nv} sites): max abs err {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < 5e-5


def test_batched_256_instances_50_sweeps_fixed_point(instance_set):
    """After 50 sweeps every instance's messages are a fixed point of the oracle's
    update map to the residual the library reports (plus FP32 rounding), sampled
    on the messages out of the highest-degree site of each instance."""
    inst = instance_set
    caches = [itnbp.initialize_cache(psi) for _, _, psi in inst]
    res = itnbp.update_batched_(caches, maxiter=50, precision="f32")
    worst = 0.0
    for i, ((kind, chi, psi), c, r) in enumerate(zip(inst, caches, res)):
        g = psi.graph
        og, opsi = oracle_graph(g), oracle_psi(psi)
        msgs = to_index(g, c.messages())
        u = max(range(g.nv), key=lambda x: (len(og.incident[x]), -x))
        gap = 0.0
        for v in og.neighbors(u):
            new = O.updated_message(og, opsi, msgs, u, v)
            gap = max(gap, float(np.linalg.norm(new - msgs[(u, v)])))
        slack = gap - 2 * r["max_residual"]
        worst = max(worst, slack)
        assert gap <= 2 * r["max_residual"] + 2e-5, (i, kind, chi, gap, r["max_residual"])
    print(f"256 instances: worst oracle fixed-point gap beyond the reported residual {worst:.2e}", flush=True)


@pytest.fixture(scope="module")
def peps32():
    """32x32 open PEPS, chi = 32, d = 2, N(0, 1/chi) rounded to BF16."""
    from itnbp.cache import bf16_round
    g = itnbp.named_grid((32, 32))
    rng = np.random.default_rng(1234)
    tensors = {}
    for v in g.vertices:
        shape = (2,) + (32,) * g.degree(v)
        x = rng.standard_normal(shape, dtype=np.float32) * np.float32(1 / np.sqrt(32))
        tensors[v] = bf16_round(x)[0].reshape(shape).astype(np.float64)
    psi = itnbp.ITensorNetwork(g, tensors)
    return g, psi, itnbp.initialize_cache(psi, storage="bf16")


@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("bf16x3", 1e-6), ("f16x2", 2e-6), ("f32", 1e-3)])
def test_peps32x32_chi32_one_sweep_sampled_messages(peps32, precision, tol):
    g, psi, bpc = peps32
    bpc.ctx.check(bpc.lib.itnbp_init_identity_messages(bpc.h))
    itnbp.update_(bpc, maxiter=1, edge_sequence="parallel", precision=precision)
    path = bpc.last_update["path"]
    print(f"32x32 chi=32 {precision}: path {path}", flush=True)
    assert "chi32" in path
    got = bpc.messages()
    og = oracle_graph(g)
    rng = np.random.default_rng(8)
    corners = [v for v in g.vertices if g.degree(v) == 2]
    edges3 = [v for v in g.vertices if g.degree(v) == 3]
    bulk = [v for v in g.vertices if g.degree(v) == 4]
    srcs = corners + [edges3[i] for i in rng.choice(len(edges3), 8, replace=False)] + \
        [bulk[i] for i in rng.choice(len(bulk), 8, replace=False)]
    worst = 0.0
    for u in srcs:
        iu = g.index[u]
        # the incoming messages of u are identities after identity initialisation
        msgs = {(g.index[w], iu): np.eye(32) for w in g.incident(u)}
        full = [None] * g.nv
        full[iu] = np.asarray(psi[u], dtype=np.float64)
        for w in g.incident(u)[:3] if g.degree(u) == 4 else g.incident(u):
            ref = O.updated_message(og, full, msgs, iu, g.index[w])
            worst = max(worst, float(np.abs(got[(u, w)] - ref).max()))
    print(f"{precision}: max abs err over sampled messages {worst:.2e}", flush=True)
    assert worst < tol
