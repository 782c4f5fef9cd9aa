"""GPU tests of API semantics that the parity tests leave open: edge groups,
non-finite inputs, the site order of two-site operators and gates, and the
refusal of an apply call whose kept rank cannot fit the bond."""
import numpy as np
import pytest

import itnbp
from itnbp import _lib as L
from helpers import oracle_graph, oracle_psi, to_index
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


def make(dims, chi, seed, dtype=np.float64):
    g = itnbp.named_grid(dims)
    return g, itnbp.random_tensornetwork(np.random.default_rng(seed), dtype, g, link_space=chi)


def converged(psi):
    return itnbp.update(itnbp.initialize_cache(psi), maxiter=200, res_tol=1e-13, edge_sequence="parallel")


def test_edge_groups_follow_the_grouped_update():
    """Groups of two edges (abstractbeliefpropagationcache.jl:294-308): each group
    sequential from the sweep's starting messages, groups merged; the stopping
    measure divides by the number of groups."""
    g, psi = make((3, 3), 3, 9)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    seq = itnbp.default_edge_sequence(itnbp.initialize_cache(psi))
    groups = [seq[i:i + 2] for i in range(0, len(seq), 2)]
    ogroups = [[(g.index[a], g.index[b]) for (a, b) in grp] for grp in groups]
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=4, edge_sequence=groups)
    ref, it, _ = O.bp_update_groups(og, opsi, O.identity_messages(og, opsi), ogroups, maxiter=4)
    got = to_index(g, bpc.messages())
    assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-12
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=50, tol=1e-9, edge_sequence=groups)
    _, it_ref, md = O.bp_update_groups(og, opsi, O.identity_messages(og, opsi), ogroups, maxiter=50, tol=1e-9)
    assert bpc.last_update["iterations"] == it_ref
    assert bpc.last_update["mean_diff"] == pytest.approx(md, rel=1e-6, abs=1e-15)
    # directed edges given as two-element lists are a sequential schedule, not groups
    lst = [[a, b] for (a, b) in seq]
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=2, edge_sequence=lst)
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=2, schedule="sequential",
                               sequence=[(g.index[a], g.index[b]) for (a, b) in seq])
    got = to_index(g, bpc.messages())
    assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-12


def with_nan(psi, v):
    bad = psi.copy()
    t = np.array(bad[v], copy=True)
    t.flat[3] = np.nan
    bad[v] = t
    return bad


@pytest.mark.parametrize("schedule", ["parallel", None])
def test_nonfinite_site_tensor_fails_update(schedule):
    g, psi = make((4, 4), 4, 1)
    bpc = itnbp.initialize_cache(with_nan(psi, (2, 2)))
    with pytest.raises(itnbp.ITNBPError) as ei:
        itnbp.update_(bpc, maxiter=3, edge_sequence=schedule)
    assert ei.value.code == L.ERR_NONFINITE


def test_nonfinite_site_tensor_fails_batched_update():
    g, psi = make((4, 4), 4, 1)
    caches = [itnbp.initialize_cache(psi), itnbp.initialize_cache(with_nan(psi, (3, 2)))]
    with pytest.raises(itnbp.ITNBPError) as ei:
        itnbp.update_batched_(caches, maxiter=3)
    assert ei.value.code == L.ERR_NONFINITE
    assert "instance 1" in str(ei.value)


def test_nonfinite_site_tensor_fails_apply():
    g, psi = make((3, 3), 2, 5)
    bpc = converged(psi)
    bpc[(2, 2)] = with_nan(psi, (2, 2))[(2, 2)]
    G = O.expm_sym(O.ising_bond_hamiltonian(2.5, 4, 3), -0.05)
    with pytest.raises(itnbp.ITNBPError) as ei:
        itnbp.apply_gates_(bpc, [itnbp.Gate(G, ((2, 2), (2, 1)))], maxdim=2)
    assert ei.value.code == L.ERR_NONFINITE


def test_asymmetric_two_site_operator_site_order():
    """<X_u Z_v> with O[(s_u', s_v'), (s_u, s_v)], s_u fastest, on bonds named in
    both orientations, against the oracle's pair RDMs."""
    g, psi = make((3, 3), 3, 4)
    bpc = converged(psi)
    og, opsi = oracle_graph(g), oracle_psi(psi)
    msgs = to_index(g, bpc.messages())
    XZ = O.two_site(O.X, O.Z)
    vals = {}
    for (u, v) in [((1, 1), (2, 1)), ((2, 1), (1, 1)), ((2, 2), (2, 3)), ((2, 3), (2, 2))]:
        got = itnbp.expect2_on_cache(bpc, XZ, edges=[(u, v)])[(u, v)]
        ref = O.expect2(XZ, O.rdm2(og, opsi, msgs, g.index[u], g.index[v]))
        assert got == pytest.approx(ref, abs=1e-11)
        vals[(u, v)] = got
    assert abs(vals[((1, 1), (2, 1))] - vals[((2, 1), (1, 1))]) > 1e-3


def test_asymmetric_gate_on_a_reversed_bond_is_exact():
    """An untruncated non-symmetric gate named (v, u) against the graph's edge
    (u, v): psi_v' psi_u' = G (psi_v psi_u) with G acting on (s_v, s_u)."""
    g = itnbp.named_grid((2, 2))
    psi = itnbp.random_tensornetwork(np.random.default_rng(1), np.float64, g, link_space=[8, 2, 2, 2])
    bpc = converged(psi)
    u, v = (1, 1), (2, 1)
    assert g.edges[g.edge_id(u, v)] == (u, v)
    G = O.expm_sym(O.two_site(O.X, O.Z) + 0.7 * O.two_site(O.Z, O.I2), -0.3)
    new = itnbp.apply(itnbp.Gate(G, (v, u)), psi, envs=bpc, maxdim=8)
    og = oracle_graph(g)

    def pair(p, a, b):
        ta, tb = p[a], p[b]
        t = np.tensordot(ta, tb, axes=([og.axis(a, b)], [og.axis(b, a)]))
        return np.moveaxis(t, ta.ndim - 1, 1)

    iu, iv = g.index[u], g.index[v]
    ref = pair(oracle_psi(psi), iv, iu)
    rhs = np.einsum("klij,ij...->kl...", G.reshape(2, 2, 2, 2, order="F"), ref)
    lhs = pair(oracle_psi(new), iv, iu)
    np.testing.assert_allclose(lhs, rhs, atol=1e-11 * np.abs(rhs).max())
