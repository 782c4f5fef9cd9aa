"""Pins for the CPU oracle (no GPU).

The reference holds no numeric fixtures for this path (its BP tests are
relational), so each oracle function is pinned by something that does not
restate its formula: exact contraction of the whole network, tree exactness of
one forest-cover sweep, product states, a single brute-force einsum written
from the index definition, LAPACK identities, and the closed-form invariant of
an untruncated simple update.
"""
import itertools
import string

import numpy as np
import pytest

from oracle import bp_oracle as O


def brute_force_message(g, psi, msgs, u, v):
    """One einsum over every index of the message u->v, labels from the definition
    T[b_v, b'_v] = sum psi[s, b] conj(psi[s, b']) prod_{w != v} M_{w->u}[b_w, b'_w]."""
    letters = iter(string.ascii_letters)
    s = next(letters)
    ket = [s]
    bra = [s]
    ops, terms = [], []
    out = None
    for (w, _) in g.incident[u]:
        b, bb = next(letters), next(letters)
        ket.append(b)
        bra.append(bb)
        if w == v:
            out = b + bb
        else:
            terms.append(b + bb)
            ops.append(msgs[(w, u)])
    spec = ",".join(["".join(ket), "".join(bra)] + terms) + "->" + out
    return np.einsum(spec, psi[u], psi[u].conj(), *ops)


@pytest.fixture
def grid33():
    g = O.square_lattice(3, 3)
    rng = np.random.default_rng(1234)
    psi = O.random_peps(g, 2, 2, rng)
    return g, psi


def test_lattice_shapes():
    g = O.square_lattice(4, 4)
    assert g.nv == 16 and len(g.edges) == 24
    g = O.square_lattice(16, 16)
    assert len(g.edges) == 480 and len(g.directed_edges()) == 960
    g = O.square_lattice(32, 32)
    assert len(g.edges) == 1984
    hh = O.heavy_hex_lattice(4, 3)
    degs = [len(x) for x in hh.incident]
    assert max(degs) <= 3 and min(degs) >= 1


def test_forest_cover_sequence_covers_each_directed_edge_once():
    for g in (O.square_lattice(4, 4), O.square_lattice(5, 3), O.comb_tree(3, 3),
              O.heavy_hex_lattice(3, 3)):
        seq = O.forest_cover_sequence(g)
        assert len(seq) == 2 * len(g.edges)
        assert set(seq) == set(g.directed_edges())


def test_message_matches_brute_force_einsum(grid33):
    g, psi = grid33
    rng = np.random.default_rng(7)
    msgs = {}
    for (u, v) in g.directed_edges():
        a = rng.standard_normal((2, 2))
        msgs[(u, v)] = a  # deliberately non-symmetric: pins the (ket, bra) index order
    for (u, v) in g.directed_edges():
        ref = brute_force_message(g, psi, msgs, u, v)
        got = O.updated_message(g, psi, msgs, u, v, normalize=False)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
        nrm = O.updated_message(g, psi, msgs, u, v)
        np.testing.assert_allclose(nrm, ref / np.linalg.norm(ref), rtol=1e-12)


def test_message_diff_closed_forms():
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    assert abs(O.message_diff(a, a)) < 1e-15
    assert abs(O.message_diff(a, -3 * a)) < 1e-15
    b = np.array([[0.0, 1.0], [0.0, 0.0]])
    c = np.array([[1.0, 0.0], [0.0, 0.0]])
    assert O.message_diff(b, c) == pytest.approx(1.0)
    # cos^2 = 1/2 for 45 degrees
    assert O.message_diff(c, b + c) == pytest.approx(0.5)


@pytest.mark.parametrize("g", [O.path_graph(6), O.comb_tree(3, 3)])
def test_tree_one_sweep_is_exact(g):
    """BP on a tree from an empty cache with the forest-cover order is exact after one
    sweep (test_expect.jl:13-20, test_inner.jl:12-38, test_forms.jl:59-70)."""
    rng = np.random.default_rng(11)
    psi = O.random_peps(g, 2, 3, rng)
    msgs = O.initialize_messages(g, psi)
    assert msgs == {}
    msgs, it, _, _ = O.bp_update(g, psi, msgs, maxiter=1)
    state = O.exact_state(g, psi)
    for v in range(g.nv):
        bp = O.expect1(O.SZ, O.rdm1(g, psi, msgs, v))
        ex = O.expect1(O.SZ, O.exact_rdm1(state, v))
        assert bp == pytest.approx(ex, rel=1e-11, abs=1e-13)
    for (u, v) in g.edges:
        rb = O.rdm2(g, psi, msgs, u, v)
        re = O.exact_rdm2(state, u, v)
        np.testing.assert_allclose(rb / np.trace(rb), re / np.trace(re), atol=1e-12)
    # log Z_BP = log <psi|psi> on a tree (test_inner.jl:12-38)
    assert O.logscalar(g, psi, msgs) == pytest.approx(np.log(O.exact_norm2(state)), rel=1e-11)


def test_product_state_bp_is_exact_on_a_loopy_grid():
    """Product states make BP exact (test_expect.jl:41-50)."""
    g = O.square_lattice(3, 3)
    psi = []
    for v in range(g.nv):
        shape = (2,) + (1,) * len(g.incident[v])
        t = np.zeros(shape)
        t[(v % 2,) + (0,) * len(g.incident[v])] = 1.0
        psi.append(t)
    msgs = O.initialize_messages(g, psi)
    msgs, _, _, _ = O.bp_update(g, psi, msgs, maxiter=5, tol=1e-14)
    for v in range(g.nv):
        assert O.expect1(O.SZ, O.rdm1(g, psi, msgs, v)) == pytest.approx(0.5 if v % 2 == 0 else -0.5)


def test_loopy_fixed_point_and_schedule_independence(grid33):
    """update(maxiter=25, tol=eps) leaves every message at a fixed point:
    message_diff(updated, current) < 10 eps (test_belief_propagation.jl:49-53);
    the sequential and Jacobi schedules reach the same fixed point."""
    g, psi = grid33
    m0 = O.initialize_messages(g, psi)
    ms, _, _, _ = O.bp_update(g, psi, m0, maxiter=25, tol=O.EPS64, schedule="sequential")
    for (u, v) in g.directed_edges():
        new = O.updated_message(g, psi, ms, u, v)
        assert O.message_diff(new, ms[(u, v)]) < 10 * O.EPS64
    # message_diff ~ angle^2 cannot certify 1e-10; run both schedules a fixed, large
    # number of sweeps and compare the messages themselves
    ms, _, _, _ = O.bp_update(g, psi, m0, maxiter=150, schedule="sequential")
    mj, _, _, _ = O.bp_update(g, psi, m0, maxiter=300, schedule="parallel")
    for (u, v) in g.directed_edges():
        np.testing.assert_allclose(ms[(u, v)], mj[(u, v)], atol=1e-12)
        # PSD, symmetric, unit norm
        m = ms[(u, v)]
        np.testing.assert_allclose(m, m.T, atol=1e-12)
        assert np.min(np.linalg.eigvalsh(m)) > -1e-12
        assert np.linalg.norm(m) == pytest.approx(1.0)


def test_rdm2_hermitian_psd(grid33):
    g, psi = grid33
    msgs, _, _, _ = O.bp_update(g, psi, O.initialize_messages(g, psi), maxiter=25, tol=1e-14)
    u, v = 4, 7  # (2,2)-(2,3) of test_belief_propagation.jl:70
    rho = O.rdm2(g, psi, msgs, u, v)
    rho = rho / np.trace(rho)
    assert rho.shape == (4, 4)
    np.testing.assert_allclose(rho, rho.T, atol=1e-13)
    assert np.min(np.linalg.eigvalsh(rho)) >= -1e-14


def test_rdm2_contains_rdm1(grid33):
    """Tracing out one site of the BP pair RDM leaves the one-site BP RDM computed with
    the message from the partner in place of the partner itself - this holds at the
    BP fixed point."""
    g, psi = grid33
    msgs, _, _, _ = O.bp_update(g, psi, O.initialize_messages(g, psi), maxiter=400, tol=1e-15)
    u, v = 4, 5
    rho = O.rdm2(g, psi, msgs, u, v).reshape(2, 2, 2, 2, order="F")  # [su, sv, su', sv']
    ru = np.einsum("abcb->ac", rho)
    r1 = O.rdm1(g, psi, msgs, u)
    np.testing.assert_allclose(ru / np.trace(ru), r1 / np.trace(r1), atol=1e-10)


def test_zero_tensor_gives_zero_scalar():
    """test_belief_propagation.jl:93-99."""
    g = O.path_graph(3)
    rng = np.random.default_rng(1234)
    psi = O.random_peps(g, 2, 2, rng)
    psi[0] = 0 * psi[0]
    msgs, _, _, _ = O.bp_update(g, psi, O.initialize_messages(g, psi), maxiter=1)
    assert np.exp(O.logscalar(g, psi, msgs)) == 0.0


def test_normalize_bp_makes_every_scalar_one():
    """test_normalize.jl:59-66."""
    g = O.square_lattice(3, 2)
    rng = np.random.default_rng(5)
    psi = O.random_peps(g, 2, 2, rng)
    msgs, _, _, _ = O.bp_update(g, psi, O.initialize_messages(g, psi), maxiter=100, tol=1e-15)
    psi2, msgs2 = O.normalize_bp(g, psi, msgs)
    msgs3, _, _, _ = O.bp_update(g, psi2, msgs2, maxiter=1)
    for (u, v) in g.edges:
        assert O.edge_scalar(msgs2, u, v) == pytest.approx(1.0)
    for u in range(g.nv):
        assert O.vertex_scalar(g, psi2, msgs2, u) == pytest.approx(1.0)
    assert O.logscalar(g, psi2, msgs2) == pytest.approx(0.0, abs=1e-12)


def test_map_eigvals_identities():
    """test_map_eigvals.jl:25-33."""
    rng = np.random.default_rng(3)
    for n in (2, 3, 5, 10):
        A = rng.standard_normal((n, n))
        P = A @ A.T + 0.1 * np.eye(n)
        s = O.map_eigvals(np.sqrt, P)
        si = O.map_eigvals(lambda x: 1 / np.sqrt(x), P)
        np.testing.assert_allclose(s @ s.T, P, atol=1e-12)
        np.testing.assert_allclose(si @ P @ si, np.eye(n), atol=1e-10)
        np.testing.assert_allclose(si @ s, np.eye(n), atol=1e-10)


def test_truncation_rule():
    p = np.array([4.0, 3.0, 2.0, 1.0])
    assert O.truncation_keep(p, None, 2) == (2, pytest.approx(0.3))
    assert O.truncation_keep(p, 0.1, None) == (3, pytest.approx(0.1))
    assert O.truncation_keep(p, 0.09, None)[0] == 4


def _pair(g, psi, u, v):
    """psi_u . psi_v over their shared bond -> [su, sv, ext_u..., ext_v...]."""
    au, av = g.axis(u, v), g.axis(v, u)
    t = np.tensordot(psi[u], psi[v], axes=([au], [av]))
    # axes: su, ext_u..., sv, ext_v...
    ku = psi[u].ndim - 2
    return np.moveaxis(t, ku + 1, 1)


def test_simple_update_closed_form_without_truncation(grid33):
    """Full-rank envs and no truncation: the √M and M^{-1/2} factors cancel and
    psi_u' psi_v' = G (psi_u psi_v) exactly, for any env (apply.jl:33-95)."""
    g, psi = grid33
    msgs, _, _, _ = O.bp_update(g, psi, O.initialize_messages(g, psi), maxiter=50, tol=1e-14)
    rng = np.random.default_rng(9)
    G = rng.standard_normal((4, 4))
    for (u, v) in [(4, 5), (0, 1), (1, 4)]:
        new, s, err = O.apply_gate(G, g, psi, msgs, u, v)
        assert err == 0.0
        lhs = _pair(g, new, u, v)
        ref = _pair(g, psi, u, v)
        G4 = G.reshape(2, 2, 2, 2, order="F")
        rhs = np.einsum("klij,ij...->kl...", G4, ref)
        np.testing.assert_allclose(lhs, rhs, atol=1e-10 * np.abs(rhs).max())


def test_simple_update_truncation_matches_lapack(grid33):
    g, psi = grid33
    msgs, _, _, _ = O.bp_update(g, psi, O.initialize_messages(g, psi), maxiter=50, tol=1e-14)
    G = O.expm_sym(O.ising_bond_hamiltonian(2.5, 3, 4), -0.05)
    u, v = 1, 4
    full, s_full, _ = O.apply_gate(G, g, psi, msgs, u, v)
    new, s, err = O.apply_gate(G, g, psi, msgs, u, v, maxdim=2, normalize=True)
    assert len(s) == 2 and err > 0
    np.testing.assert_allclose(s, s_full[:2], rtol=1e-12)
    assert err == pytest.approx(np.sum(s_full[2:] ** 2) / np.sum(s_full ** 2))
    assert np.linalg.norm(new[u]) == pytest.approx(1.0)
    assert new[u].shape == psi[u].shape


def test_exact_state_small_closed_form():
    """Two sites, one bond of dim 1: the state is the outer product."""
    g = O.path_graph(2)
    a = np.array([[1.0], [2.0]])
    b = np.array([[3.0], [-1.0]])
    st = O.exact_state(g, [a, b])
    np.testing.assert_allclose(st, np.outer(a[:, 0], b[:, 0]))
    rho = O.exact_rdm2(st, 0, 1)
    vec = st.reshape(-1, order="F")
    np.testing.assert_allclose(rho, np.outer(vec, vec))


def test_planted_mistake_is_caught(grid33):
    """A transposed incoming message (a plausible index mix-up) changes the message
    when messages are not symmetric: the brute-force pin detects it."""
    g, psi = grid33
    rng = np.random.default_rng(2)
    msgs = {e: rng.standard_normal((2, 2)) for e in g.directed_edges()}
    bad = {e: m.T for e, m in msgs.items()}
    u, v = 4, 5
    ref = brute_force_message(g, psi, msgs, u, v)
    wrong = O.updated_message(g, psi, bad, u, v, normalize=False)
    assert np.abs(wrong - ref).max() > 1e-3


def test_grad_oracle_energy_matches_numpy_oracle_and_finite_differences():
    """The torch restatement of the unrolled BP energy equals the NumPy oracle's,
    and its autograd gradient matches central finite differences (step 1e-6)."""
    from oracle import grad_oracle as GO
    g = O.square_lattice(3, 3)
    rng = np.random.default_rng(4)
    psi = O.random_peps(g, 2, 2, rng)
    h = 2.5
    ops1 = [h * O.X for _ in range(g.nv)]
    ops2 = [O.two_site(O.Z, O.Z) for _ in g.edges]
    E, grads = GO.energy_and_grad(g, psi, 10, ops1, ops2)
    msgs, _, _, _ = O.bp_update(g, psi, O.identity_messages(g, psi), maxiter=10, schedule="parallel")
    assert E == pytest.approx(O.bp_ising_energy(g, psi, msgs, h), rel=1e-12)
    rng = np.random.default_rng(9)
    for _ in range(3):
        dirs = [rng.standard_normal(p.shape) for p in psi]
        eps = 1e-6

        def energy(shift):
            q = [p + shift * d for p, d in zip(psi, dirs)]
            m, _, _, _ = O.bp_update(g, q, O.identity_messages(g, q), maxiter=10, schedule="parallel")
            return O.bp_ising_energy(g, q, m, h)

        fd = (energy(eps) - energy(-eps)) / (2 * eps)
        an = sum(float(np.sum(gr * d)) for gr, d in zip(grads, dirs))
        assert an == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_grouped_schedule_reduces_to_the_jacobi_and_sequential_sweeps(grid33):
    """Edge groups (abstractbeliefpropagationcache.jl:294-308): singleton groups are
    the Jacobi sweep, and one group holding the whole forest-cover order is the
    sequential sweep.  A grouped update that did not restart every group from the
    sweep's starting messages fails the first identity."""
    g, psi = grid33
    m0 = O.initialize_messages(g, psi)
    seq = O.forest_cover_sequence(g)
    mj, _, _, _ = O.bp_update(g, psi, m0, maxiter=3, schedule="parallel", sequence=seq)
    mg, it, _ = O.bp_update_groups(g, psi, m0, [[e] for e in seq], maxiter=3)
    assert it == 3
    for e in seq:
        np.testing.assert_allclose(mg[e], mj[e], rtol=0, atol=1e-14)
    ms, _, _, _ = O.bp_update(g, psi, m0, maxiter=3, schedule="sequential", sequence=seq)
    mg1, _, _ = O.bp_update_groups(g, psi, m0, [list(seq)], maxiter=3)
    for e in seq:
        np.testing.assert_allclose(mg1[e], ms[e], rtol=0, atol=1e-14)
    # two-edge groups differ from both (a path the Jacobi-flattened schedule misses)
    pairs = [seq[i:i + 2] for i in range(0, len(seq), 2)]
    mp, _, _ = O.bp_update_groups(g, psi, m0, pairs, maxiter=1)
    mj1, _, _, _ = O.bp_update(g, psi, m0, maxiter=1, schedule="parallel", sequence=seq)
    assert max(np.abs(mp[e] - mj1[e]).max() for e in seq) > 1e-6


def test_expm_sym_matches_scipy():
    """The gate builder exp(t H) of a real symmetric H against scipy.linalg.expm
    (Pade scaling and squaring, an independent method)."""
    from scipy.linalg import expm
    rng = np.random.default_rng(11)
    for n in (2, 4, 8):
        a = rng.standard_normal((n, n))
        H = a + a.T
        for t in (-0.05, 0.3):
            np.testing.assert_allclose(O.expm_sym(H, t), expm(t * H), rtol=1e-12, atol=1e-13)
    H = O.ising_bond_hamiltonian(2.5, 4, 3)
    np.testing.assert_allclose(O.expm_sym(H, -0.05), expm(-0.05 * H), rtol=1e-12, atol=1e-13)


def test_torch_simple_update_matches_the_oracle_and_finite_differences(grid33):
    """The torch restatement of simple_update_bp (the checker of the gate
    application's adjoint) equals the NumPy oracle, and its autograd VJP equals
    central finite differences of the NumPy oracle along random directions of
    psi_u, psi_v, the gate and every environment message."""
    import torch
    from oracle import grad_oracle as GO
    g, psi = grid33
    rng = np.random.default_rng(71)
    msgs, _, _, _ = O.bp_update(g, psi, O.identity_messages(g, psi), maxiter=30, schedule="parallel")
    T = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64, requires_grad=True)
    for (u, v), maxdim, norm in [((4, 5), None, False), ((0, 1), 2, True), ((1, 4), 2, False)]:
        envs_u, envs_v = O.envs_for_bond(g, msgs, u, v)
        G = O.expm_sym(O.ising_bond_hamiltonian(1.0, len(g.incident[u]), len(g.incident[v])), -0.3)
        ax_u, ax_v = g.axis(u, v), g.axis(v, u)
        pu, pv, _, _ = O.simple_update_bp(G, psi[u], psi[v], ax_u, ax_v, envs_u, envs_v, maxdim=maxdim,
                                          normalize=norm)
        cu, cv = rng.standard_normal(pu.shape), rng.standard_normal(pv.shape)
        args = dict(G=T(G), psi_u=T(psi[u]), psi_v=T(psi[v]))
        eu = {a: T(m) for a, m in envs_u.items()}
        ev = {a: T(m) for a, m in envs_v.items()}
        tu, tv = GO.simple_update_bp_torch(args["G"], args["psi_u"], args["psi_v"], ax_u, ax_v, eu, ev,
                                           maxdim=maxdim, normalize=norm)
        np.testing.assert_allclose(tu.detach().numpy(), pu, atol=1e-12)
        L = (tu * torch.tensor(cu)).sum() + (tv * torch.tensor(cv)).sum()
        L.backward()

        def loss(G_, pu_, pv_, eu_, ev_):
            a, b, _, _ = O.simple_update_bp(G_, pu_, pv_, ax_u, ax_v, eu_, ev_, maxdim=maxdim, normalize=norm)
            return float(np.sum(a * cu) + np.sum(b * cv))

        base = [G, psi[u], psi[v], dict(envs_u), dict(envs_v)]
        grads = [args["G"].grad, args["psi_u"].grad, args["psi_v"].grad, {a: t.grad for a, t in eu.items()},
                 {a: t.grad for a, t in ev.items()}]
        h = 1e-6
        for slot in range(5):
            keys = list(base[slot].keys()) if slot >= 3 else [None]
            for key in keys:
                x0 = base[slot][key] if key is not None else base[slot]
                gr = (grads[slot][key] if key is not None else grads[slot]).numpy()
                dirn = rng.standard_normal(x0.shape)
                if slot >= 3:
                    dirn = dirn + dirn.T  # messages are symmetric
                vals = []
                for sgn in (1, -1):
                    cur = [b if not isinstance(b, dict) else dict(b) for b in base]
                    if key is None:
                        cur[slot] = x0 + sgn * h * dirn
                    else:
                        cur[slot][key] = x0 + sgn * h * dirn
                    vals.append(loss(*cur))
                fd = (vals[0] - vals[1]) / (2 * h)
                assert np.sum(gr * dirn) == pytest.approx(fd, rel=1e-5, abs=1e-7), ((u, v), slot, key)
