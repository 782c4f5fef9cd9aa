"""GPU parity of the BP sweep, environments and scalars against the CPU oracle,
through the C ABI (libitnbp.so)."""
import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, oracle_psi, oracle_sequence, to_index, to_named
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


def make(dims, chi, seed, scale=1.0, dtype=np.float64):
    g = itnbp.named_grid(dims)
    psi = itnbp.random_tensornetwork(np.random.default_rng(seed), dtype, g, link_space=chi, scale=scale)
    return g, psi


def max_msg_err(g, a, b):
    return max(np.abs(a[e] - b[e]).max() for e in a)


def test_forest_cover_sequence_matches_oracle():
    for dims in ((4, 4), (5, 3), (16, 16)):
        g, psi = make(dims, 2, 0)
        bpc = itnbp.initialize_cache(psi)
        seq = itnbp.default_edge_sequence(bpc)
        og = oracle_graph(g)
        assert oracle_sequence(g, seq) == O.forest_cover_sequence(og)


def test_sequential_4x4_chi2_messages_and_sz():
    """4x4 open PEPS, d=2, chi=2, N(0,1), normalised; sequential schedule, 30 sweeps,
    tol 1e-12; all messages and <Sz> at every site (parity 1e-10)."""
    g, psi = make((4, 4), 2, 1234)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    # normalise the state by exact contraction (normalize.jl:117-123)
    n2 = O.exact_norm2(O.exact_state(og, opsi))
    c = n2 ** (-1.0 / (2 * g.nv))
    psi = itnbp.ITensorNetwork(g, {v: psi[v] * c for v in g.vertices})
    opsi = oracle_psi(psi)
    bpc = itnbp.initialize_cache(psi)
    bpc = itnbp.update(bpc, maxiter=30, tol=1e-12)
    ref, it, md, _ = O.bp_update(og, opsi, O.initialize_messages(og, opsi), maxiter=30, tol=1e-12)
    assert bpc.last_update["iterations"] == it
    got = to_index(g, bpc.messages())
    for e in ref:
        np.testing.assert_allclose(got[e], ref[e], rtol=0, atol=1e-10)
    sz = itnbp.expect_on_cache(bpc, "Sz")
    for i, v in enumerate(g.vertices):
        r = O.expect1(O.SZ, O.rdm1(og, opsi, ref, i))
        assert sz[v] == pytest.approx(r, abs=1e-10)


@pytest.mark.parametrize("dims,chi", [((4, 4), 3), ((3, 3), 8), ((5, 4), 4)])
def test_parallel_schedule_parity(dims, chi):
    g, psi = make(dims, chi, 7)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=12, edge_sequence="parallel")
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=12, schedule="parallel")
    got = to_index(g, bpc.messages())
    for e in ref:
        np.testing.assert_allclose(got[e], ref[e], rtol=0, atol=1e-11)


def test_large_path_chi16_parity():
    """chi=16 on a 3x3 grid: the bulk vertex (k=4, 131072 elements) exceeds LDS and
    runs on the global-scratch path."""
    g, psi = make((3, 3), 16, 3, scale=0.25)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=4, edge_sequence="parallel")
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=4, schedule="parallel")
    got = to_index(g, bpc.messages())
    for e in ref:
        np.testing.assert_allclose(got[e], ref[e], rtol=0, atol=1e-11)


def test_tree_is_exact_after_one_sweep():
    g = itnbp.comb_tree(3, 3)
    psi = itnbp.random_tensornetwork(np.random.default_rng(5), np.float64, g, link_space=3)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    bpc = itnbp.initialize_cache(psi)
    assert not bpc.messages()
    bpc = itnbp.update(bpc)  # default maxiter = 1 on a tree
    state = O.exact_state(og, opsi)
    sz = itnbp.expect_on_cache(bpc, "Sz")
    for i, v in enumerate(g.vertices):
        assert sz[v] == pytest.approx(O.expect1(O.SZ, O.exact_rdm1(state, i)), abs=1e-12)
    assert itnbp.logscalar(bpc) == pytest.approx(np.log(O.exact_norm2(state)), rel=1e-12)


def test_rdm2_expect2_logscalar_and_scalars():
    g, psi = make((4, 3), 3, 11)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=40, tol=1e-14)
    ref = to_index(g, bpc.messages())
    for (a, b) in g.edges[:6]:
        r = itnbp.rdm(bpc, [a, b])
        rr = O.rdm2(og, opsi, ref, g.index[a], g.index[b])
        np.testing.assert_allclose(r, rr, rtol=1e-11, atol=1e-12 * np.abs(rr).max())
    zz = itnbp.expect2_on_cache(bpc, O.two_site(O.Z, O.Z))
    for (a, b), val in zz.items():
        rr = O.expect2(O.two_site(O.Z, O.Z), O.rdm2(og, opsi, ref, g.index[a], g.index[b]))
        assert val == pytest.approx(rr, abs=1e-11)
    vs, es = itnbp.region_scalars(bpc)
    for i in range(g.nv):
        assert vs[i] == pytest.approx(O.vertex_scalar(og, opsi, ref, i), rel=1e-12)
    for k, (a, b) in enumerate(og.edges):
        assert es[k] == pytest.approx(O.edge_scalar(ref, a, b), rel=1e-12)
    assert itnbp.logscalar(bpc) == pytest.approx(O.logscalar(og, opsi, ref), rel=1e-12)


def test_normalize_makes_scalars_one():
    g, psi = make((3, 3), 2, 4)
    cache = [None]
    cache[0] = itnbp.update(itnbp.initialize_cache(psi), maxiter=60, tol=1e-15)
    psi2 = itnbp.normalize(psi, cache=cache, update_cache=False)
    vs, es = itnbp.region_scalars(cache[0])
    np.testing.assert_allclose(vs, 1.0, rtol=1e-12)
    np.testing.assert_allclose(es, 1.0, rtol=1e-12)
    og = oracle_graph(g)
    p_ref, _ = O.normalize_bp(og, oracle_psi(psi), to_index(g, itnbp.update(itnbp.initialize_cache(psi), maxiter=60, tol=1e-15).messages()))
    for i, v in enumerate(g.vertices):
        np.testing.assert_allclose(psi2[v], p_ref[i], rtol=1e-10, atol=1e-12)


def test_fixed_point_and_message_diff():
    """test_belief_propagation.jl:49-53 on a 3x3 grid."""
    g, psi = make((3, 3), 2, 1234)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=25, tol=np.finfo(np.float64).eps)
    for (a, b) in g.directed_edges():
        assert itnbp.message_diff(itnbp.updated_message(bpc, (a, b)), bpc.message((a, b))) < 10 * np.finfo(float).eps


def test_zero_tensor_scalar_is_zero():
    g = itnbp.path_graph(3)
    psi = itnbp.random_tensornetwork(np.random.default_rng(1234), np.float64, g, link_space=2)
    psi[1] = 0 * psi[1]
    bpc = itnbp.update(itnbp.initialize_cache(psi))
    assert itnbp.scalar(bpc) == 0.0


def test_errors():
    g, psi = make((3, 3), 2, 0)
    with pytest.raises(ValueError, match="number of iterations"):
        itnbp.update(itnbp.initialize_cache(psi))
    bpc = itnbp.BeliefPropagationCache(psi)  # no messages on a loopy graph
    with pytest.raises(itnbp.ITNBPError):
        itnbp.update(bpc, maxiter=2)
    with pytest.raises(itnbp.ITNBPError):
        bpc.message((g.vertices[0], g.vertices[1]))


def test_f32_mode_close_to_f64():
    g, psi = make((6, 6), 4, 21, scale=0.5)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=30, schedule="parallel")
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=30, edge_sequence="parallel", precision="f32")
    got = to_index(g, bpc.messages())
    assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-5


def test_bf16_storage_then_f64_refinement():
    """BF16 site tensors: the low phase runs until the residual < 1e-4, then FP64 sweeps
    reach the FP64 fixed point of the BF16-rounded tensors (parity 1e-10)."""
    g, psi = make((5, 5), 4, 8, scale=0.5)
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    psi_b = bpc.tensornetwork()
    og, opsi = oracle_graph(g), oracle_psi(psi_b)
    bpc = itnbp.update(bpc, maxiter=200, res_tol=1e-12, edge_sequence="parallel", precision="bf16_f64",
                       switch_tol=1e-4)
    assert bpc.last_update["iterations_low"] >= 1
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=300, schedule="parallel")
    got = to_index(g, bpc.messages())
    assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-10


def _random_messages_small(g, psi, rng, symmetric):
    msgs = {}
    for (u, v) in g.edges:
        chi = psi[u].shape[g.axis(u, v)]
        for e in ((u, v), (v, u)):
            a = rng.standard_normal((chi, chi))
            msgs[e] = (a + a.T) / 2 if symmetric else a
    return msgs


@pytest.mark.parametrize("symmetric", [True, False])
def test_vertex_kernel_and_generic_path_with_user_messages(symmetric):
    """Jacobi sweeps from user messages on a mixed-degree graph (a 3x4 grid, chi 2..5
    per bond): symmetric ones (eigenvalues of both signs) run on the per-vertex
    eigenbasis kernel, non-symmetric ones on the generic kernels; both match the
    oracle."""
    g = itnbp.named_grid((3, 4))
    rng = np.random.default_rng(21 if symmetric else 22)
    chis = [int(c) for c in rng.integers(2, 6, size=g.ne)]
    psi = itnbp.random_tensornetwork(rng, np.float64, g, link_space=chis)
    msgs = _random_messages_small(g, psi, rng, symmetric)
    bpc = itnbp.initialize_cache(psi)
    bpc.set_messages_(msgs)
    out = itnbp.update(bpc, maxiter=3, edge_sequence="parallel")
    assert ("[vertex" in out.last_update["path"]) == symmetric
    og, opsi = oracle_graph(g), oracle_psi(psi)
    start = {(g.index[a], g.index[b]): m for (a, b), m in msgs.items()}
    ref, _, _, _ = O.bp_update(og, opsi, start, maxiter=3, schedule="parallel")
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    assert err < 1e-11, err


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_vertex_kernel_matches_generic_kernels(precision, monkeypatch):
    """The same Jacobi sweeps with and without the per-vertex kernels: a 5x5 grid at
    chi = 6 (degrees 2..4) and a 2x6 grid at chi = 16 (degrees 2..3) in LDS; the
    degree-4 sites of a 3x4 grid at chi = 12 and a 3x3 grid at chi = 16 and the
    degree-6 centre of a 3x3 triangular lattice at chi = 6, too large for LDS
    (scratch tensors in global memory, k_vertex_big)."""
    g0 = itnbp.named_grid((3, 3))
    tri = itnbp.NamedGraph(g0.vertices, list(g0.edges) + [((i, j), (i + 1, j + 1)) for i in (1, 2) for j in (1, 2)])
    assert max(tri.degree(v) for v in tri.vertices) == 6
    for g, chi in ((itnbp.named_grid((5, 5)), 6), (itnbp.named_grid((2, 6)), 16), (itnbp.named_grid((3, 4)), 12),
                   (itnbp.named_grid((3, 3)), 16), (tri, 6)):
        psi = itnbp.random_tensornetwork(np.random.default_rng(5), np.float64, g, link_space=chi)
        a = itnbp.update(itnbp.initialize_cache(psi), maxiter=6, edge_sequence="parallel", precision=precision)
        assert "[vertex" in a.last_update["path"]
        monkeypatch.setenv("ITNBP_DISABLE_VERTEX", "1")
        b = itnbp.update(itnbp.initialize_cache(psi), maxiter=6, edge_sequence="parallel", precision=precision)
        monkeypatch.delenv("ITNBP_DISABLE_VERTEX")
        assert "[vertex" not in b.last_update["path"]
        ma, mb = a.messages(), b.messages()
        tol = 1e-12 if precision == "f64" else 2e-5
        assert max(np.abs(ma[e] - mb[e]).max() for e in ma) < tol
