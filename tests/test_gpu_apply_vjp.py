"""GPU parity of the gate-application adjoint (itnbp_apply_vjp, k_apply_vjp)
against torch FP64 autograd of the torch restatement of simple_update_bp
(oracle/grad_oracle.py, itself pinned by finite differences of the NumPy
oracle), and against central finite differences of the device forward
(apply_gates_) on a 3x3, chi = 3 PEPS.  The bonds of a colour class mix bulk
sides (E >= d chi, tall) and corner sides (E < d chi, wide)."""
import numpy as np
import pytest
import torch

import itnbp
from helpers import oracle_graph, oracle_psi, to_index
from oracle import bp_oracle as O
from oracle import grad_oracle as GO

pytestmark = pytest.mark.gpu


def setup(seed, dims=(3, 3), chi=3):
    g = itnbp.named_grid(dims)
    psi = itnbp.random_tensornetwork(np.random.default_rng(seed), np.float64, g, link_space=chi)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=60, res_tol=1e-12, edge_sequence="parallel")
    return g, bpc


def gates_for(g, bonds, t=-0.3):
    return [itnbp.Gate(O.expm_sym(O.ising_bond_hamiltonian(1.0, g.degree(u), g.degree(v)), t), (u, v))
            for (u, v) in bonds]


def forward(bpc, gates, maxdim, normalize):
    c = bpc.copy()
    itnbp.apply_gates_(c, gates, maxdim=maxdim, normalize=normalize)
    new = c.tensornetwork()
    kept = [new.bond_dim(*gt.sites) for gt in gates]
    return new, kept


@pytest.mark.parametrize("maxdim,normalize", [(None, False), (2, True), (3, False)])
def test_apply_vjp_matches_torch_autograd(maxdim, normalize):
    g, bpc = setup(81)
    og = oracle_graph(g)
    opsi = oracle_psi(bpc.tensornetwork())
    msgs = to_index(g, bpc.messages())
    bonds = itnbp.edge_coloring(g)[0]
    gates = gates_for(g, bonds)
    new, kept = forward(bpc, gates, maxdim, normalize)
    rng = np.random.default_rng(82)
    dout = [(rng.standard_normal(new[u].shape), rng.standard_normal(new[v].shape)) for (u, v) in bonds]
    dpsi, dgates, dmsgs = itnbp.apply_gates_vjp(bpc, gates, kept, dout, maxdim=maxdim, normalize=normalize)
    worst = 0.0
    for i, gt in enumerate(gates):
        u, v = gt.sites
        iu, iv = g.index[u], g.index[v]
        envs_u, envs_v = O.envs_for_bond(og, msgs, iu, iv)
        T = lambda a: torch.tensor(np.asarray(a), dtype=torch.float64, requires_grad=True)
        tG, tu, tv = T(gt.matrix), T(opsi[iu]), T(opsi[iv])
        eu = {a: T(m) for a, m in envs_u.items()}
        ev = {a: T(m) for a, m in envs_v.items()}
        pu, pv = GO.simple_update_bp_torch(tG, tu, tv, og.axis(iu, iv), og.axis(iv, iu), eu, ev, maxdim=maxdim,
                                           normalize=normalize)
        np.testing.assert_allclose(new[u], pu.detach().numpy(), atol=1e-10)
        np.testing.assert_allclose(new[v], pv.detach().numpy(), atol=1e-10)
        # the kept singular values: the pullback divides by s_j and by s_j^2 - s_i^2,
        # so FP64 agreement degrades with kappa = s_max / s_min of what is kept
        _, _, sk, _ = O.simple_update_bp(gt.matrix, opsi[iu], opsi[iv], og.axis(iu, iv), og.axis(iv, iu), envs_u,
                                         envs_v, maxdim=maxdim, normalize=normalize)
        kappa = float(sk[0] / sk[-1])
        tol = max(1e-9, 1e-14 * kappa ** 2)
        L = (pu * torch.tensor(dout[i][0])).sum() + (pv * torch.tensor(dout[i][1])).sum()
        L.backward()
        pairs = [(dpsi[i][0], tu.grad), (dpsi[i][1], tv.grad), (dgates[i], tG.grad)]
        for side, x, envs in ((eu, iu, envs_u), (ev, iv, envs_v)):
            for a, t in side.items():
                w = og.incident[x][a - 1][0]
                pairs.append((dmsgs[(g.vertices[w], g.vertices[x])], t.grad))
        for got, ref in pairs:
            ref = ref.numpy()
            scale = max(1.0, np.abs(ref).max())
            err = np.abs(got - ref).max() / scale
            worst = max(worst, err)
            assert err < tol, (i, (u, v), err, kappa)
        print(f"gate {(u, v)}: kappa {kappa:.3g}", flush=True)
    print(f"maxdim={maxdim} normalize={normalize}: kept {kept}, worst relative error {worst:.2e}", flush=True)


# the untruncated case keeps every singular value of Theta; at t = -0.2 the
# smallest are ~1e-5 of the largest (kappa ~ 1.5e5, the pullback's FP64 error
# ~ eps kappa^2 ~ 1e-5), at t = -1 kappa <= 600
@pytest.mark.parametrize("maxdim,normalize,t", [(2, True, -0.2), (None, False, -1.0)])
def test_apply_vjp_matches_finite_differences_of_the_device_forward(maxdim, normalize, t):
    """L = sum_i <c_i, apply_gates_(psi, msgs)_i>: the directional derivative
    along random directions of every site tensor and environment message from
    the VJP against central differences (steps 1e-6 and 5e-7, Richardson
    extrapolated) of the device forward, truncated and untruncated (the bond
    grows to the rank of Theta)."""
    g, bpc = setup(83)
    bonds = itnbp.edge_coloring(g)[1]
    gates = gates_for(g, bonds, t=t)
    new, kept = forward(bpc, gates, maxdim, normalize)
    rng = np.random.default_rng(84)
    cots = [(rng.standard_normal(new[u].shape), rng.standard_normal(new[v].shape)) for (u, v) in bonds]
    dpsi, dgates, dmsgs = itnbp.apply_gates_vjp(bpc, gates, kept, cots, maxdim=maxdim, normalize=normalize)
    psi0 = bpc.tensornetwork()
    msgs0 = bpc.messages()

    def loss(psi_mod, msg_mod):
        c = itnbp.initialize_cache(psi_mod)
        for e, m in msg_mod.items():
            c.set_message_(e, m)
        itnbp.apply_gates_(c, gates, maxdim=maxdim, normalize=normalize)
        out = c.tensornetwork()
        return sum(float(np.sum(out[u] * cu) + np.sum(out[v] * cv)) for (u, v), (cu, cv) in zip(bonds, cots))

    h = 1e-6
    for trial in range(3):
        dpsi_dir = {v: rng.standard_normal(psi0[v].shape) for v in g.vertices}
        dmsg_dir = {}
        for e in dmsgs:
            r = rng.standard_normal(msgs0[e].shape)
            dmsg_dir[e] = r + r.T
        pred = 0.0
        for i, (u, v) in enumerate(bonds):
            pred += float(np.sum(dpsi[i][0] * dpsi_dir[u]) + np.sum(dpsi[i][1] * dpsi_dir[v]))
        pred += sum(float(np.sum(dmsgs[e] * dmsg_dir[e])) for e in dmsgs)
        def central(step):
            vals = []
            for sgn in (1, -1):
                pm = psi0.copy()
                for v in g.vertices:
                    pm[v] = psi0[v] + sgn * step * dpsi_dir[v]
                mm = dict(msgs0)
                for e in dmsg_dir:
                    mm[e] = msgs0[e] + sgn * step * dmsg_dir[e]
                vals.append(loss(pm, mm))
            return (vals[0] - vals[1]) / (2 * step)

        # Richardson extrapolation of two central differences removes their
        # O(h^2) term: the untruncated case keeps environment eigenvalues and
        # singular values near 1e-5 of the largest, where a step of 1e-6 is
        # not yet in the linear regime
        f1, f2 = central(h), central(h / 2)
        fd = (4 * f2 - f1) / 3
        print(f"direction {trial}: vjp {pred:.10f} finite differences {f1:.10f} (h), {f2:.10f} (h/2), "
              f"extrapolated {fd:.10f}", flush=True)
        assert pred == pytest.approx(fd, rel=1e-6, abs=1e-8)


@pytest.mark.parametrize("checkpoint", [1, 3])
def test_gradient_of_a_tebd_step(checkpoint):
    """One TEBD step: T Jacobi sweeps from identity messages, then a gate round
    on one colour class with those messages as environments.  The gradient of
    L = sum_i <c_i, psi'_i> with respect to every site tensor (the apply adjoint
    for the direct dependence, the sweep adjoint for the one through the
    messages) against central finite differences of the whole device forward."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(91), np.float64, g, link_space=2)
    T = 6
    bonds = itnbp.edge_coloring(g)[0]
    gates = gates_for(g, bonds, t=-0.25)
    maxdim, normalize = 2, True

    def step(p):
        c = itnbp.update(itnbp.initialize_cache(p), maxiter=T, edge_sequence="parallel")
        fwd = c.copy()
        itnbp.apply_gates_(fwd, gates, maxdim=maxdim, normalize=normalize)
        return c, fwd.tensornetwork()

    bpc, new = step(psi)
    kept = [new.bond_dim(*gt.sites) for gt in gates]
    rng = np.random.default_rng(92)
    cots = [(rng.standard_normal(new[u].shape), rng.standard_normal(new[v].shape)) for (u, v) in bonds]
    dpsi, _, dmsgs = itnbp.apply_gates_vjp(bpc, gates, kept, cots, maxdim=maxdim, normalize=normalize)
    grad = itnbp.sweeps_vjp(psi, T, dmsgs, checkpoint=checkpoint)
    for i, (u, v) in enumerate(bonds):
        grad[u] = grad[u] + dpsi[i][0]
        grad[v] = grad[v] + dpsi[i][1]

    def loss(p):
        _, out = step(p)
        return sum(float(np.sum(out[u] * cu) + np.sum(out[v] * cv)) for (u, v), (cu, cv) in zip(bonds, cots))

    h = 1e-6
    for trial in range(3):
        d = {v: rng.standard_normal(psi[v].shape) for v in g.vertices}
        pred = sum(float(np.sum(grad[v] * d[v])) for v in g.vertices)
        vals = []
        for sgn in (1, -1):
            pm = psi.copy()
            for v in g.vertices:
                pm[v] = psi[v] + sgn * h * d[v]
            vals.append(loss(pm))
        fd = (vals[0] - vals[1]) / (2 * h)
        print(f"TEBD step, direction {trial}: gradient {pred:.10f} finite differences {fd:.10f}", flush=True)
        assert pred == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_apply_vjp_refuses_bonds_past_the_lds_kernel():
    """The adjoint covers the forward's LDS kernel only: a bond with an external
    bond of dimension 12 (the forward runs it on the global-memory variant) is
    refused with ITNBP_ERR_UNSUPPORTED, and the cache is left as it was."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(91), np.float64, g, link_space=12)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=30, tol=1e-10, edge_sequence="parallel")
    before = bpc.tensornetwork()
    G = O.expm_sym(O.ising_bond_hamiltonian(2.5, 4, 3), -0.1)
    gates = [itnbp.Gate(G, ((2, 2), (2, 3)))]
    new = bpc.copy()
    errs = itnbp.apply_gates_(new, gates, maxdim=12)  # the forward runs
    assert len(errs) == 1
    dout = [(np.ones_like(new.tensornetwork()[(2, 2)]), np.ones_like(new.tensornetwork()[(2, 3)]))]
    with pytest.raises(RuntimeError, match="external bond dimensions <= 8"):
        itnbp.apply_gates_vjp(bpc, gates, [12], dout, maxdim=12)
    after = bpc.tensornetwork()
    for v in g.vertices:
        np.testing.assert_array_equal(after[v], before[v])
