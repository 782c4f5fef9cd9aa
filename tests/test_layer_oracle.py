"""Pins of the two-layer (bra != ket, real or complex) BP oracle
(oracle/layer_oracle.py) by facts it does not restate: on trees one
forest-cover sweep gives the exact <phi|psi> and <phi|A|psi> of the whole
network (brute-force contraction); with phi = psi real it reproduces the pinned
norm-network oracle; per-site phases of a complex state cancel; the messages of
a complex norm network stay Hermitian."""
import numpy as np
import pytest

from oracle import bp_oracle as O
from oracle import layer_oracle as LO


def rand(rng, shape, cplx):
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


def random_net(g, chi, rng, cplx, d=2, scale=0.5):
    out = []
    for v in range(g.nv):
        shape = (d,) + tuple(chi if np.isscalar(chi) else chi[e] for (_, e) in g.incident[v])
        out.append(scale * rand(rng, shape, cplx))
    return out


def random_tree(rng, n):
    return O.Graph(n, [(int(rng.integers(0, v)), v) for v in range(1, n)])


TREES = [("path5", lambda r: O.path_graph(5)), ("comb3x2", lambda r: O.comb_tree(3, 2)),
         ("random7", lambda r: random_tree(r, 7))]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("name,make", TREES, ids=[t[0] for t in TREES])
def test_tree_inner_is_exact(name, make, cplx):
    rng = np.random.default_rng(3)
    g = make(rng)
    ket = random_net(g, 2, rng, cplx)
    bra = random_net(g, 3, rng, cplx)  # different bond dimension in the bra layer
    msgs, _, _ = LO.bp_update2(g, ket, bra, {}, maxiter=1)
    assert LO.logscalar2(g, ket, bra, msgs) == pytest.approx(np.log(LO.exact_inner(g, ket, bra)), rel=1e-11,
                                                             abs=1e-11)


@pytest.mark.parametrize("cplx", [False, True])
def test_tree_operator_inner_is_exact(cplx):
    rng = np.random.default_rng(5)
    g = O.comb_tree(3, 2)
    ket = random_net(g, 2, rng, cplx)
    bra = random_net(g, 2, rng, cplx)
    A = []
    for v in range(g.nv):
        shape = (2, 2) + tuple(2 for _ in g.incident[v])
        A.append(0.7 * rand(rng, shape, cplx))
    ketA = LO.fold_operator(g, ket, A)
    msgs, _, _ = LO.bp_update2(g, ketA, bra, {}, maxiter=1)
    exact = LO.exact_operator_inner(g, bra, A, ket)
    assert LO.logscalar2(g, ketA, bra, msgs) == pytest.approx(np.log(exact), rel=1e-11, abs=1e-11)
    # a product operator (bond dimension 1) is the local operators' action
    A1 = [np.asarray(a[(slice(None), slice(None)) + (0,) * (a.ndim - 2)]).reshape((2, 2) + (1,) * (a.ndim - 2))
          for a in A]
    ket1 = LO.fold_operator(g, ket, A1)
    local = [np.tensordot(A1[v].reshape(2, 2), ket[v], axes=(1, 0)) for v in range(g.nv)]
    for v in range(g.nv):
        np.testing.assert_allclose(ket1[v], local[v], atol=1e-14)


def test_equal_layers_reproduce_the_norm_network_oracle():
    rng = np.random.default_rng(7)
    g = O.square_lattice(3, 3)
    psi = random_net(g, 3, rng, False)
    ref, it, _, _ = O.bp_update(g, psi, O.identity_messages(g, psi), maxiter=12, schedule="parallel")
    got, it2, _ = LO.bp_update2(g, psi, psi, LO.identity_messages2(g, psi), maxiter=12, schedule="parallel")
    for e in ref:
        np.testing.assert_allclose(got[e], ref[e], atol=1e-13)
    assert LO.logscalar2(g, psi, psi, got) == pytest.approx(O.logscalar(g, psi, ref), rel=1e-12)
    for u in range(g.nv):
        np.testing.assert_allclose(LO.rdm1_2(g, psi, psi, got, u), O.rdm1(g, psi, ref, u), atol=1e-13)


def test_site_phases_cancel_in_a_complex_norm_network():
    rng = np.random.default_rng(9)
    g = O.square_lattice(3, 3)
    psi = random_net(g, 2, rng, False)
    phases = np.exp(1j * rng.uniform(0, 2 * np.pi, g.nv))
    psic = [phases[v] * psi[v] for v in range(g.nv)]
    # Jacobi sweeps: the normalisation does not damp the anti-Hermitian part of a
    # complex message, and sequential sweeps amplify its rounding noise (by about
    # 10x per sweep here) into a non-Hermitian fixed point; 12 Jacobi sweeps keep
    # it at the rounding level
    ref, _, _, _ = O.bp_update(g, psi, O.identity_messages(g, psi), maxiter=12, schedule="parallel")
    got, _, _ = LO.bp_update2(g, psic, psic, LO.identity_messages2(g, psic), maxiter=12, schedule="parallel")
    for e in ref:
        np.testing.assert_allclose(got[e], ref[e], atol=1e-13)
    assert LO.logscalar2(g, psic, psic, got) == pytest.approx(O.logscalar(g, psi, ref), rel=1e-12, abs=1e-12)


def test_complex_norm_network_messages_stay_hermitian():
    rng = np.random.default_rng(11)
    g = O.square_lattice(3, 3)
    psi = random_net(g, 3, rng, True)
    msgs, _, _ = LO.bp_update2(g, psi, psi, LO.identity_messages2(g, psi), maxiter=6, schedule="parallel")
    for e, m in msgs.items():
        np.testing.assert_allclose(m, m.conj().T, atol=1e-13)
    # and the BP norm of a complex state is real and positive at the fixed point
    ls = LO.logscalar2(g, psi, psi, msgs)
    assert abs(np.imag(ls)) < 1e-10


def test_planted_conjugation_mistake_is_caught():
    """Leaving the bra unconjugated changes a complex inner product on a tree."""
    rng = np.random.default_rng(13)
    g = O.path_graph(4)
    ket = random_net(g, 2, rng, True)
    bra = random_net(g, 2, rng, True)
    msgs, _, _ = LO.bp_update2(g, ket, bra, {}, maxiter=1)
    good = LO.logscalar2(g, ket, bra, msgs)
    wrong = LO.logscalar2(g, ket, [np.conj(b) for b in bra],
                          LO.bp_update2(g, ket, [np.conj(b) for b in bra], {}, maxiter=1)[0])
    assert abs(good - np.log(LO.exact_inner(g, ket, bra))) < 1e-10
    assert abs(wrong - np.log(LO.exact_inner(g, ket, bra))) > 1e-3


# ---- several link indices per edge -----------------------------------------
from oracle import multilink_oracle as ML  # noqa: E402


def multilink_net(g, link_dims, rng, d=2, scale=0.6):
    psi = []
    for v in range(g.nv):
        shape = [d]
        for (_, e) in g.incident[v]:
            shape += list(link_dims[e])
        psi.append(scale * rng.standard_normal(shape))
    return psi


def test_multilink_tree_is_exact():
    rng = np.random.default_rng(31)
    g = O.comb_tree(3, 2)
    link_dims = [(2, 3) if e % 2 == 0 else (2,) for e in range(len(g.edges))]
    psi = multilink_net(g, link_dims, rng)
    msgs = ML.bp_update(g, link_dims, psi, {}, maxiter=1)
    assert ML.logscalar(g, link_dims, psi, msgs) == pytest.approx(np.log(ML.exact_norm2(g, link_dims, psi)),
                                                                  rel=1e-11)


def test_multilink_with_one_link_per_edge_is_the_norm_network_oracle():
    rng = np.random.default_rng(32)
    g = O.square_lattice(3, 3)
    link_dims = [(3,)] * len(g.edges)
    psi = multilink_net(g, link_dims, rng)
    ref, _, _, _ = O.bp_update(g, psi, O.identity_messages(g, psi), maxiter=10, schedule="parallel")
    got = ML.bp_update(g, link_dims, psi, ML.identity_messages(g, link_dims), maxiter=10, schedule="parallel")
    for e in ref:
        np.testing.assert_allclose(got[e], ref[e], atol=1e-13)


def test_merged_bonds_reproduce_the_multilink_messages():
    """Host merge (itnbp.merge_link_indices, first link fastest) + the one-bond
    oracle against the per-link oracle on a loopy grid."""
    import itnbp
    from helpers import oracle_graph
    rng = np.random.default_rng(33)
    g = itnbp.named_grid((3, 3))
    og = oracle_graph(g)
    link_dims = {e: ((2, 2) if i % 3 else (3, 2)) for i, e in enumerate(g.edges)}
    ld = [link_dims[e] for e in g.edges]
    psi_l = multilink_net(og, ld, rng)
    merged = itnbp.merge_link_indices(g, {v: psi_l[g.index[v]] for v in g.vertices}, link_dims)
    psi_m = [np.asarray(merged[v]) for v in g.vertices]
    ref = ML.bp_update(og, ld, psi_l, ML.identity_messages(og, ld), maxiter=8, schedule="parallel")
    got, _, _, _ = O.bp_update(og, psi_m, O.identity_messages(og, psi_m), maxiter=8, schedule="parallel")
    for (u, w), m in ref.items():
        dims = ld[og.edge_id(u, w)]
        np.testing.assert_allclose(itnbp.split_message(got[(u, w)], dims), m, atol=1e-13)
    assert O.logscalar(og, psi_m, got) == pytest.approx(ML.logscalar(og, ld, psi_l, ref).real, rel=1e-12)
