"""GPU parity of the BP-gauged gate application (simple update, batched FP64
Jacobi SVD) against the CPU oracle."""
import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, oracle_psi, to_index
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


def converged_cache(psi, sweeps=200):
    return itnbp.update(itnbp.initialize_cache(psi), maxiter=sweeps, res_tol=1e-13, edge_sequence="parallel")


def pair(g, psi, u, v):
    au, av = g.axis(u, v), g.axis(v, u)
    t = np.tensordot(psi[u], psi[v], axes=([au], [av]))
    return np.moveaxis(t, psi[u].ndim - 1, 1)


def test_untruncated_gate_is_exact_closed_form():
    """With full-rank envs and no truncation, psi_u' psi_v' = G psi_u psi_v for any
    env (apply.jl:33-95): the bond (1,1)-(2,1) has chi = 8 >= rank(Theta) = 4."""
    g = itnbp.named_grid((2, 2))
    psi = itnbp.random_tensornetwork(np.random.default_rng(1), np.float64, g, link_space=[8, 2, 2, 2])
    bpc = converged_cache(psi)
    G = np.random.default_rng(2).standard_normal((4, 4))
    u, v = (1, 1), (2, 1)
    svals = []
    new = itnbp.apply(itnbp.Gate(G, (u, v)), psi, envs=bpc, maxdim=8,
                      callback=lambda singular_values, truncation_error: svals.append((singular_values, truncation_error)))
    og = oracle_graph(g)
    ref = pair(og, oracle_psi(psi), g.index[u], g.index[v])
    rhs = np.einsum("klij,ij...->kl...", G.reshape(2, 2, 2, 2, order="F"), ref)
    lhs = pair(og, oracle_psi(new), g.index[u], g.index[v])
    np.testing.assert_allclose(lhs, rhs, atol=1e-11 * np.abs(rhs).max())
    msgs = to_index(g, bpc.messages())
    _, s_ref, err_ref = O.apply_gate(G, og, oracle_psi(psi), msgs, g.index[u], g.index[v], maxdim=8)
    s, err = svals[0]
    np.testing.assert_allclose(s, s_ref, rtol=1e-12)
    assert err == 0.0 and err_ref == 0.0


def test_simple_update_round_matches_oracle():
    """One BP-gauged round of exp(-0.05 H_b), H_b = ZZ + 2.5 (X I/deg_u + I X/deg_v),
    on every bond of a 4x4 chi=4 PEPS, truncated back to chi (config-2 pattern)."""
    g = itnbp.named_grid((4, 4))
    psi = itnbp.random_tensornetwork(np.random.default_rng(3), np.float64, g, link_space=4)
    bpc = converged_cache(psi)
    og = oracle_graph(g)
    opsi = oracle_psi(psi)
    msgs = to_index(g, bpc.messages())
    deg = {v: g.degree(v) for v in g.vertices}

    def gate_named(u, v):
        return O.expm_sym(O.ising_bond_hamiltonian(2.5, deg[u], deg[v]), -0.05)

    errs = itnbp.simple_update_round_(bpc, gate_named, maxdim=4, normalize=True)
    new = bpc.tensornetwork()
    odeg = [len(og.incident[i]) for i in range(og.nv)]
    ref_psi, ref_errs, _ = O.simple_update_round(
        og, opsi, msgs, lambda a, b: O.expm_sym(O.ising_bond_hamiltonian(2.5, odeg[a], odeg[b]), -0.05),
        maxdim=4, normalize=True)
    for i, v in enumerate(g.vertices):
        np.testing.assert_allclose(new[v], ref_psi[i], atol=1e-10)
    for (a, b), e in errs.items():
        assert e == pytest.approx(ref_errs[(g.index[a], g.index[b])], rel=1e-8, abs=1e-15)
        assert e > 0


def test_truncated_random_gate_fidelity_and_errors():
    """test_apply.jl:12-66 pattern on a 2x2 chi=2 grid: truncation error > 0,
    fidelity with the exact gate application > 0, non-product env throws."""
    g = itnbp.named_grid((2, 2))
    psi = itnbp.random_tensornetwork(np.random.default_rng(1234), np.float64, g, link_space=2)
    bpc = converged_cache(psi)
    u, v = (2, 2), (1, 2)
    og = oracle_graph(g)
    rng = np.random.default_rng(5)
    for _ in range(3):
        Q, _ = np.linalg.qr(rng.standard_normal((4, 4)))
        te = []
        new = itnbp.apply(itnbp.Gate(Q, (u, v)), psi, envs=itnbp.environment(bpc, [u, v]), maxdim=2,
                          normalize=True, callback=lambda singular_values, truncation_error: te.append(truncation_error))
        assert te[0] > 0
        exact = O.exact_state(og, oracle_psi(psi))
        G4 = Q.reshape(2, 2, 2, 2, order="F")
        iu, iv = g.index[u], g.index[v]
        ex = np.moveaxis(np.tensordot(G4, np.moveaxis(exact, [iu, iv], [0, 1]), axes=([2, 3], [0, 1])), [0, 1], [iu, iv])
        st = O.exact_state(og, oracle_psi(new))
        f = np.vdot(st, ex) / np.sqrt(np.vdot(st, st) * np.vdot(ex, ex))
        assert abs(f) ** 2 > 0
    with pytest.raises(RuntimeError, match="product environment"):
        bad = dict(itnbp.environment(bpc, [u, v]))
        bad[(u, v)] = np.eye(2)
        itnbp.apply(itnbp.Gate(np.eye(4), (u, v)), psi, envs=bad, maxdim=2)
    with pytest.raises(ValueError, match="neighbors"):
        itnbp.apply(itnbp.Gate(np.eye(4), ((1, 1), (2, 2))), psi, envs=bpc, maxdim=2)


def test_tebd_with_bp_regauging_matches_oracle():
    """Three steps of exp(-0.1 H_b) rounds on a 3x3 chi=3 PEPS, each preceded by 8
    Jacobi sweeps warm-started from the previous messages (fixed sweep counts so
    both sides do identical work); tensors after the last round match the
    oracle's loop and the BP energy goes down."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(11), np.float64, g, link_space=3)
    deg = {v: g.degree(v) for v in g.vertices}
    bpc = itnbp.initialize_cache(psi)
    og = oracle_graph(g)
    odeg = [len(og.incident[i]) for i in range(og.nv)]
    opsi = oracle_psi(psi)
    omsgs = O.identity_messages(og, opsi)
    oe = [O.bp_ising_energy(og, opsi, O.bp_update(og, opsi, omsgs, maxiter=30, schedule="parallel")[0], 2.5)]
    hist = itnbp.tebd_bp_(bpc, lambda u, v: O.expm_sym(O.ising_bond_hamiltonian(2.5, deg[u], deg[v]), -0.1), 3,
                          maxdim=3, bp_maxiter=8, bp_tol=None)
    assert len(hist) == 3 and all(h >= 0 for h in hist)
    for _ in range(3):
        omsgs, _, _, _ = O.bp_update(og, opsi, omsgs, maxiter=8, schedule="parallel")
        opsi, _, _ = O.simple_update_round(
            og, opsi, omsgs, lambda a, b: O.expm_sym(O.ising_bond_hamiltonian(2.5, odeg[a], odeg[b]), -0.1),
            maxdim=3, normalize=True)
        oe.append(O.bp_ising_energy(og, opsi, O.bp_update(og, opsi, omsgs, maxiter=30, schedule="parallel")[0], 2.5))
    new = bpc.tensornetwork()
    for i, v in enumerate(g.vertices):
        np.testing.assert_allclose(new[v], opsi[i], atol=1e-9)
    assert oe[-1] < oe[0]


def test_one_site_gates_run_on_the_device():
    """One-site gates (apply.jl:110-116) mixed with a two-site gate in one call:
    psi_v <- O psi_v (normalised) on the device, against the host contraction;
    the two-site gate against the oracle."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(12), np.float64, g, link_space=3)
    bpc = converged_cache(psi)
    og = oracle_graph(g)
    msgs = to_index(g, bpc.messages())
    rng = np.random.default_rng(13)
    ops = {v: rng.standard_normal((2, 2)) for v in [(1, 1), (3, 3), (2, 1)]}
    G = O.expm_sym(O.ising_bond_hamiltonian(2.5, 4, 3), -0.1)
    gates = [itnbp.Gate(m, (v,)) for v, m in ops.items()] + [itnbp.Gate(G, ((2, 2), (2, 3)))]
    errs = itnbp.apply_gates_(bpc, gates, maxdim=3, normalize=True)
    assert errs[:3] == [0.0, 0.0, 0.0]
    new = bpc.tensornetwork()
    for v, m in ops.items():
        t = np.tensordot(m, psi[v], axes=([1], [0]))
        np.testing.assert_allclose(new[v], t / np.linalg.norm(t), atol=1e-14)
    ref, _, err = O.apply_gate(G, og, oracle_psi(psi), msgs, g.index[(2, 2)], g.index[(2, 3)], maxdim=3,
                               normalize=True)
    for v in [(2, 2), (2, 3)]:
        np.testing.assert_allclose(new[v], ref[g.index[v]], atol=1e-10)
    assert errs[3] == pytest.approx(err, rel=1e-8)
    # the functional form
    out = itnbp.apply(itnbp.Gate(ops[(1, 1)], ((1, 1),)), psi)
    t = np.tensordot(ops[(1, 1)], psi[(1, 1)], axes=([1], [0]))
    np.testing.assert_allclose(out[(1, 1)], t, atol=1e-14)


def test_bond_takes_the_kept_dimension():
    """apply.jl:81-93: the new bond has the kept dimension.  Without maxdim or
    cutoff a chi = 2 bond grows to the full rank of Theta (8); with a cutoff a
    bond shrinks.  Tensors, kept ranks and truncation errors against the oracle,
    which builds the new bond from the kept singular values; the messages of a
    resized bond become identities."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(21), np.float64, g, link_space=2)
    bpc = converged_cache(psi)
    og = oracle_graph(g)
    msgs = to_index(g, bpc.messages())
    G = O.expm_sym(O.two_site(O.X, O.Z) + 0.5 * O.ising_bond_hamiltonian(2.5, 4, 3), -0.4)
    u, v = (2, 2), (2, 3)
    kept = []
    new = itnbp.apply(itnbp.Gate(G, (u, v)), psi, envs=bpc,
                      callback=lambda singular_values, truncation_error: kept.append(len(singular_values)))
    ref, s_ref, err_ref = O.apply_gate(G, og, oracle_psi(psi), msgs, g.index[u], g.index[v])
    assert kept == [len(s_ref)] and len(s_ref) == 8
    for x in (u, v):
        assert new[x].shape == ref[g.index[x]].shape
        np.testing.assert_allclose(new[x], ref[g.index[x]], atol=1e-10)
    # in place on a cache, with a cutoff that shrinks the bond
    c2 = bpc.copy()
    errs = itnbp.apply_gates_(c2, [itnbp.Gate(G, ((1, 1), (2, 1)))], cutoff=1e-3)
    ref2, s2, err2 = O.apply_gate(G, og, oracle_psi(psi), msgs, g.index[(1, 1)], g.index[(2, 1)], cutoff=1e-3)
    t2 = c2.tensornetwork()
    chi_new = t2.bond_dim((1, 1), (2, 1))
    assert chi_new == len(s2)
    for x in ((1, 1), (2, 1)):
        np.testing.assert_allclose(t2[x], ref2[g.index[x]], atol=1e-10)
    assert errs[0] == pytest.approx(err2, rel=1e-8)
    if chi_new != 2:
        np.testing.assert_array_equal(c2.message(((1, 1), (2, 1))), np.eye(chi_new))
    # maxdim = 1 shrinks the bond to one; the resized cache still runs BP
    c3 = bpc.copy()
    errs3 = itnbp.apply_gates_(c3, [itnbp.Gate(G, ((1, 1), (2, 1)))], maxdim=1)
    ref3, s3, err3 = O.apply_gate(G, og, oracle_psi(psi), msgs, g.index[(1, 1)], g.index[(2, 1)], maxdim=1)
    t3 = c3.tensornetwork()
    assert t3.bond_dim((1, 1), (2, 1)) == 1 == len(s3)
    for x in ((1, 1), (2, 1)):
        np.testing.assert_allclose(t3[x], ref3[g.index[x]], atol=1e-10)
    assert errs3[0] == pytest.approx(err3, rel=1e-8) and errs3[0] > 0
    itnbp.update_(c3, maxiter=5, edge_sequence="parallel")


@pytest.mark.parametrize("chi", [16, 32])
def test_large_bonds_run_on_the_global_memory_kernel(chi):
    """Bonds past the LDS kernel's limits (external bonds above 8, Theta up to
    128 x 128) run on its global-memory variant: on a 3x3 chi = 16 or 32 PEPS a
    bulk bond (the centre's side has E = chi^3 external rows against n = 2 chi,
    the tall QR; at chi = 32 Theta is 128 x 128) and a corner bond (E = chi < n,
    the wide QR), truncated to maxdim = chi, against the oracle (apply.jl:33-95):
    tensors, truncation errors and singular values."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(31), np.float64, g, link_space=chi)
    bpc = converged_cache(psi)
    og = oracle_graph(g)
    opsi = oracle_psi(psi)
    msgs = to_index(g, bpc.messages())
    deg = {v: g.degree(v) for v in g.vertices}
    bonds = [((2, 2), (2, 3)), ((1, 1), (1, 2))]
    gates = [O.expm_sym(O.ising_bond_hamiltonian(2.5, deg[u], deg[v]), -0.1) for u, v in bonds]
    svals = []
    errs = itnbp.apply_gates_(bpc, [itnbp.Gate(G, b) for G, b in zip(gates, bonds)], maxdim=chi, normalize=True,
                              callback=lambda singular_values, truncation_error: svals.append(singular_values))
    new = bpc.tensornetwork()
    for i, ((u, v), G) in enumerate(zip(bonds, gates)):
        ref, s_ref, err_ref = O.apply_gate(G, og, opsi, msgs, g.index[u], g.index[v], maxdim=chi, normalize=True)
        assert new.bond_dim(u, v) == len(s_ref)
        for x in (u, v):
            r = ref[g.index[x]]
            assert new[x].shape == r.shape
            assert np.linalg.norm(new[x] - r) <= 1e-10 * np.linalg.norm(r), x
        np.testing.assert_allclose(svals[i], s_ref, rtol=1e-10, atol=1e-13 * s_ref[0])
        assert errs[i] == pytest.approx(err_ref, rel=1e-8, abs=1e-15)


def test_mixed_small_and_large_bonds_and_a_rank_deficient_side():
    """One call with gates on both kernels on a 4x4 PEPS whose bonds at (2,2)
    have chi = 12 and all others 4: the gates next to (2,2) run on the
    global-memory variant, the far ones in LDS.  One large side is
    rank-deficient (the slices b >= 6 of (2,2)'s gate bond vanish, so X has rank
    n / 2 with E >> n: the Gram QR drops the null directions where the oracle's
    Householder QR keeps zero rows).  Tensors, singular values and truncation
    errors against the oracle."""
    g = itnbp.named_grid((4, 4))
    big_v = (2, 2)
    link = [12 if big_v in (a, b) else 4 for (a, b) in g.edges]
    psi = itnbp.random_tensornetwork(np.random.default_rng(41), np.float64, g, link_space=link)
    t = psi[big_v].copy()
    idx = [slice(None)] * t.ndim
    idx[g.axis(big_v, (2, 3))] = slice(6, None)
    t[tuple(idx)] = 0.0
    psi[big_v] = t
    bpc = converged_cache(psi)
    og = oracle_graph(g)
    opsi = oracle_psi(psi)
    msgs = to_index(g, bpc.messages())
    deg = {v: g.degree(v) for v in g.vertices}
    # global-memory variant: the first two; LDS kernel: the last two
    bonds = [((2, 2), (2, 3)), ((1, 1), (1, 2)), ((3, 4), (4, 4)), ((4, 1), (4, 2))]
    gates = [O.expm_sym(O.two_site(O.X, O.Z) + O.ising_bond_hamiltonian(2.5, deg[u], deg[v]), -0.3) for u, v in bonds]
    svals = []
    errs = itnbp.apply_gates_(bpc, [itnbp.Gate(G, b) for G, b in zip(gates, bonds)], maxdim=6, normalize=True,
                              callback=lambda singular_values, truncation_error: svals.append(singular_values))
    new = bpc.tensornetwork()
    for i, ((u, v), G) in enumerate(zip(bonds, gates)):
        ref, s_ref, err_ref = O.apply_gate(G, og, opsi, msgs, g.index[u], g.index[v], maxdim=6, normalize=True)
        assert new.bond_dim(u, v) == len(s_ref)
        for x in (u, v):
            r = ref[g.index[x]]
            assert np.linalg.norm(new[x] - r) <= 1e-10 * np.linalg.norm(r), (u, v, x)
        np.testing.assert_allclose(svals[i], s_ref, rtol=1e-10, atol=1e-13 * s_ref[0])
        assert errs[i] == pytest.approx(err_ref, rel=1e-8, abs=1e-15)


def test_large_bonds_with_physical_dimension_three():
    """The global-memory variant at d = 3 (qutrits): a 3x3 chi = 12 PEPS, a bulk
    bond and one next to a corner (E = 12 < n = 36, the wide side) with random
    symmetric gates exp(-0.2 H) (9 x 9), truncated to maxdim = 12, against the
    oracle."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(51), np.float64, g, link_space=12, phys_dim=3)
    bpc = converged_cache(psi)
    og = oracle_graph(g)
    opsi = oracle_psi(psi)
    msgs = to_index(g, bpc.messages())
    rng = np.random.default_rng(52)
    # each bond in the edge's own orientation: the deterministic bond gauge
    # (the largest entry of every new-bond slice of the first site made
    # positive) then refers to the same site in the kernel and in the oracle
    bonds = [((2, 2), (3, 2)), ((1, 2), (1, 3))]
    gates = []
    for _ in bonds:
        H = rng.standard_normal((9, 9))
        gates.append(O.expm_sym(0.5 * (H + H.T), -0.2))
    errs = itnbp.apply_gates_(bpc, [itnbp.Gate(G, b) for G, b in zip(gates, bonds)], maxdim=12, normalize=True)
    new = bpc.tensornetwork()
    for i, ((u, v), G) in enumerate(zip(bonds, gates)):
        ref, s_ref, err_ref = O.apply_gate(G, og, opsi, msgs, g.index[u], g.index[v], maxdim=12, normalize=True)
        assert new.bond_dim(u, v) == len(s_ref)
        for x in (u, v):
            r = ref[g.index[x]]
            assert np.linalg.norm(new[x] - r) <= 1e-10 * np.linalg.norm(r), (u, v, x)
        assert errs[i] == pytest.approx(err_ref, rel=1e-8, abs=1e-15)
