"""GPU parity of the hand-written BP adjoint against FP64 autograd and central
finite differences (transverse-field Ising energy, J = 1, h = 2.5)."""
import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, oracle_psi
from oracle import bp_oracle as O
from oracle import grad_oracle as GO

pytestmark = pytest.mark.gpu

H = 2.5


def ising_ops(g):
    ops1 = {v: H * O.X for v in g.vertices}
    ops2 = {e: O.two_site(O.Z, O.Z) for e in g.edges}
    return ops1, ops2


@pytest.mark.parametrize("dims,chi,sweeps,ck", [((3, 3), 2, 6, 1), ((4, 3), 3, 10, 4)])
def test_gradient_matches_autograd(dims, chi, sweeps, ck):
    g = itnbp.named_grid(dims)
    psi = itnbp.random_tensornetwork(np.random.default_rng(7), np.float64, g, link_space=chi)
    ops1, ops2 = ising_ops(g)
    E, grad = itnbp.energy_and_gradient(psi, sweeps, ops1, ops2, checkpoint=ck)
    og = oracle_graph(g)
    Er, gr = GO.energy_and_grad(og, oracle_psi(psi), sweeps, [ops1[v] for v in g.vertices],
                                [ops2[e] for e in g.edges])
    assert E == pytest.approx(Er, rel=1e-12)
    num = sum(np.sum((grad[v] - gr[i]) ** 2) for i, v in enumerate(g.vertices))
    den = sum(np.sum(gr[i] ** 2) for i in range(g.nv))
    assert np.sqrt(num / den) < 1e-10


def test_gradient_8x8_chi4_40_sweeps_vs_finite_differences():
    """8x8 square PEPS, d=2, chi=4, N(0,1): gradient of the BP Ising energy through 40
    unrolled sweeps (checkpoint every 8) vs central finite differences, step 1e-6,
    along random directions (relative agreement 1e-6)."""
    g = itnbp.named_grid((8, 8))
    psi = itnbp.random_tensornetwork(np.random.default_rng(5), np.float64, g, link_space=4)
    ops1, ops2 = ising_ops(g)
    E, grad = itnbp.energy_and_gradient(psi, 40, ops1, ops2, checkpoint=8)
    E1, _ = itnbp.energy_and_gradient(psi, 40, ops1, ops2, checkpoint=1)
    assert E == E1
    rng = np.random.default_rng(11)
    eps = 1e-6
    for _ in range(3):
        d = {v: rng.standard_normal(psi[v].shape) for v in g.vertices}
        plus = itnbp.ITensorNetwork(g, {v: psi[v] + eps * d[v] for v in g.vertices})
        minus = itnbp.ITensorNetwork(g, {v: psi[v] - eps * d[v] for v in g.vertices})
        Ep, _ = itnbp.energy_and_gradient(plus, 40, ops1, ops2)
        Em, _ = itnbp.energy_and_gradient(minus, 40, ops1, ops2)
        fd = (Ep - Em) / (2 * eps)
        an = sum(float(np.sum(grad[v] * d[v])) for v in g.vertices)
        assert an == pytest.approx(fd, rel=1e-6)
