"""GPU parity of the two-layer BP kernels (k_layer.hip) against the two-layer
oracle: inner(phi, psi; alg="bp") and inner(phi, A, psi; alg="bp") on trees
(exact), bra != ket on a loopy grid, complex states, and the real norm network
against the one-layer kernels."""
import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, to_index
from oracle import bp_oracle as O
from oracle import layer_oracle as LO

pytestmark = pytest.mark.gpu


def rand_net(rng, g, chi, cplx, d=2, scale=0.5):
    t = {}
    for v in g.vertices:
        shape = (d,) + tuple(chi for _ in g.incident(v))
        a = rng.standard_normal(shape)
        t[v] = scale * (a + 1j * rng.standard_normal(shape) if cplx else a)
    return itnbp.ITensorNetwork(g, t)


def rand_op(rng, g, chi, cplx, d=2, scale=0.7):
    """Operator network tensors A_v[s', s, bonds] (a dict by vertex)."""
    t = {}
    for v in g.vertices:
        shape = (d, d) + tuple(chi for _ in g.incident(v))
        a = rng.standard_normal(shape)
        t[v] = scale * (a + 1j * rng.standard_normal(shape) if cplx else a)
    return t


def as_list(net, g=None):
    g = g or net.graph
    return [np.asarray(net[v]) for v in g.vertices]


def random_tree(rng, n):
    return itnbp.NamedGraph(list(range(n)), [(int(rng.integers(0, v)), v) for v in range(1, n)])


TREES = [("path6", lambda r: itnbp.path_graph(6)), ("comb4x3", lambda r: itnbp.comb_tree(4, 3)),
         ("random9", lambda r: random_tree(r, 9))]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("name,make", TREES, ids=[t[0] for t in TREES])
def test_inner_on_trees_is_exact(name, make, cplx):
    """inner(x, y; alg="bp") (inner.jl:147-160) on a tree: one forest-cover sweep
    from no messages (maxiter defaults to 1 on trees), bra and ket of different
    bond dimensions; equals the whole-network contraction (test/test_inner.jl:12-38)."""
    rng = np.random.default_rng(21)
    g = make(rng)
    x = rand_net(rng, g, 3, cplx)
    y = rand_net(rng, g, 2, cplx)
    og = oracle_graph(g)
    exact = LO.exact_inner(og, as_list(y), as_list(x))
    got = itnbp.inner(x, y, alg="bp")
    assert complex(got) == pytest.approx(exact, rel=1e-11)
    assert complex(np.exp(itnbp.loginner(x, y, alg="bp"))) == pytest.approx(exact, rel=1e-11)


@pytest.mark.parametrize("cplx", [False, True])
def test_inner_with_an_operator_network_on_a_tree_is_exact(cplx):
    """inner(x, A, y; alg="bp") with A a tree operator network of bond dimension 2
    (test/test_inner.jl:40-48): the operator folded into the ket on the device
    (merged link indices) and one BP sweep equal the exact <x|A|y>."""
    rng = np.random.default_rng(22)
    g = itnbp.comb_tree(3, 3)
    x = rand_net(rng, g, 2, cplx)
    y = rand_net(rng, g, 2, cplx)
    A = rand_op(rng, g, 2, cplx)
    og = oracle_graph(g)
    exact = LO.exact_operator_inner(og, as_list(x), as_list(A, g), as_list(y))
    got = itnbp.inner(x, A, y, alg="bp")
    assert complex(got) == pytest.approx(exact, rel=1e-11)
    # the folded ket on the device equals the oracle's fold
    c = itnbp.inner_network_cache(x, A, y)
    folded = c.folded_ket()
    ref = LO.fold_operator(og, as_list(y), as_list(A, g))
    for i, v in enumerate(g.vertices):
        np.testing.assert_allclose(folded[v], ref[i], atol=1e-13)


@pytest.mark.parametrize("schedule", ["parallel", "forest_cover"])
def test_bra_and_ket_differ_on_a_loopy_grid(schedule):
    """<phi|psi> on a 3x3 grid from delta messages: every message after 8 sweeps,
    the logscalar and the site matrices R[s_ket, s_bra] against the oracle."""
    rng = np.random.default_rng(23)
    g = itnbp.named_grid((3, 3))
    phi = rand_net(rng, g, 3, False)
    psi = rand_net(rng, g, 3, False)
    c = itnbp.inner_network_cache(phi, psi, messages="identity")
    c.update_(maxiter=8, edge_sequence=None if schedule == "forest_cover" else "parallel")
    og = oracle_graph(g)
    ket, bra = as_list(psi), as_list(phi)
    ref, _, _ = LO.bp_update2(og, ket, bra, LO.identity_messages2(og, ket), maxiter=8,
                              schedule="sequential" if schedule == "forest_cover" else "parallel")
    got = to_index(g, c.messages())
    assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-11
    # <phi|psi> may be negative: compare the scalars (the logs differ by 2 pi i)
    assert complex(c.scalar()) == pytest.approx(complex(np.exp(LO.logscalar2(og, ket, bra, ref))), rel=1e-10)
    for v in [(1, 1), (2, 2)]:
        np.testing.assert_allclose(c.rdm1(v), LO.rdm1_2(og, ket, bra, ref, g.index[v]), atol=1e-12)


def test_unset_messages_are_an_error_on_loopy_graphs():
    """A LinearFormNetwork cache has no default message (initialize_cache.jl:10-12):
    on a loopy graph the first update reads unset messages and fails."""
    rng = np.random.default_rng(24)
    g = itnbp.named_grid((2, 3))
    c = itnbp.inner_network_cache(rand_net(rng, g, 2, False), rand_net(rng, g, 3, False))
    with pytest.raises(itnbp.ITNBPError) as ei:
        c.update_(maxiter=3)
    assert ei.value.code == itnbp._lib.ERR_STATE
    with pytest.raises(ValueError, match="number of iterations"):
        c.update_()


def test_complex_state_bp_expect_and_logscalar():
    """initialize_cache of a complex state (test/test_belief_propagation.jl:18-20):
    BP on its norm network by the two-layer kernels; messages after 10 Jacobi
    sweeps, <Sz>, <Sx> and the logscalar against the oracle."""
    rng = np.random.default_rng(25)
    g = itnbp.named_grid((3, 3))
    psi = rand_net(rng, g, 2, True)
    bpc = itnbp.initialize_cache(psi)
    assert isinstance(bpc, itnbp.LayeredBPCache)
    bpc = itnbp.update(bpc, maxiter=10, edge_sequence="parallel")
    og = oracle_graph(g)
    k = as_list(psi)
    ref, _, _ = LO.bp_update2(og, k, k, LO.identity_messages2(og, k), maxiter=10, schedule="parallel")
    got = to_index(g, bpc.messages())
    assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-12
    for e, m in got.items():
        np.testing.assert_allclose(m, m.conj().T, atol=1e-12)
    sz = itnbp.expect_on_cache(bpc, "Sz")
    for v in g.vertices:
        R = LO.rdm1_2(og, k, k, ref, g.index[v])
        assert sz[v] == pytest.approx(np.trace(O.SZ @ R) / np.trace(R), abs=1e-12)
    assert complex(itnbp.logscalar(bpc)) == pytest.approx(LO.logscalar2(og, k, k, ref), abs=1e-11)


def test_real_norm_network_agrees_with_the_one_layer_kernels():
    """bra = ket = psi real: the two-layer kernels and the norm-network kernels
    (k_contract / k_vertex) give the same messages and logscalar."""
    rng = np.random.default_rng(26)
    g = itnbp.named_grid((4, 4))
    psi = itnbp.random_tensornetwork(rng, np.float64, g, link_space=3)
    a = itnbp.update(itnbp.initialize_cache(psi), maxiter=20, edge_sequence="parallel")
    b = itnbp.norm_cache(psi).update_(maxiter=20, edge_sequence="parallel")
    ma, mb = a.messages(), b.messages()
    assert max(np.abs(ma[e] - mb[e]).max() for e in ma) < 1e-12
    assert b.logscalar() == pytest.approx(itnbp.logscalar(a), rel=1e-11)
    assert b.last_update["path"].startswith("layered-f64")


def test_several_link_indices_per_edge():
    """Edges with two link indices (formnetworks/quadraticformnetwork.jl:115-119):
    merged into one bond at the host boundary (first link fastest), BP on the
    device from delta messages; every message, split back into per-link indices,
    against the per-link oracle, and on a tree the exact norm."""
    from oracle import multilink_oracle as ML
    rng = np.random.default_rng(27)
    for g, sweeps in [(itnbp.named_grid((3, 4)), 12), (itnbp.comb_tree(3, 3), 1)]:
        og = oracle_graph(g)
        link_dims = {e: ((2, 3) if i % 2 else (2, 2)) for i, e in enumerate(g.edges)}
        ld = [link_dims[e] for e in g.edges]
        tens = {}
        for v in g.vertices:
            shape = [2]
            for w in g.incident(v):
                shape += list(link_dims[(v, w)] if (v, w) in link_dims else link_dims[(w, v)])
            tens[v] = 0.6 * rng.standard_normal(shape)
        psi_l = [tens[v] for v in g.vertices]
        merged = itnbp.merge_link_indices(g, tens, link_dims)
        bpc = itnbp.initialize_cache(merged)
        if g.is_tree():
            bpc = itnbp.update(bpc)
            ref = ML.bp_update(og, ld, psi_l, {}, maxiter=1)
            assert itnbp.scalar(bpc) == pytest.approx(ML.exact_norm2(og, ld, psi_l), rel=1e-11)
        else:
            bpc = itnbp.update(bpc, maxiter=sweeps, edge_sequence="parallel")
            ref = ML.bp_update(og, ld, psi_l, ML.identity_messages(og, ld), maxiter=sweeps, schedule="parallel")
        got = to_index(g, bpc.messages())
        for (u, w), m in ref.items():
            np.testing.assert_allclose(itnbp.split_message(got[(u, w)], ld[og.edge_id(u, w)]), m, atol=1e-12)
