"""GPU test of itnbp_apply_gates_noalloc (the apply entry point that allocates no
device memory) against itnbp_apply_gates, through the C ABI: two calls in a
row on identical caches, the second with fewer gates, must give identical
truncation errors, singular values, kept ranks and site tensors."""
import numpy as np
import pytest

import itnbp
from itnbp import _lib as L
from itnbp.apply import edge_coloring
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


def call(bpc, name, edges, gates, maxdim, max_keep=8):
    g = bpc.graph
    n = len(edges)
    ids = np.array([g.edge_id(u, v) for (u, v) in edges], dtype=np.int32)
    mats = np.ascontiguousarray(np.concatenate([m.reshape(-1, order="F") for m in gates]))
    te = np.full(n, -1.0)
    sv = np.full(n * max_keep, -1.0)
    kept = np.full(n, -1, dtype=np.int32)
    rc = getattr(bpc.lib, name)(bpc.h, n, L.iptr(ids), L.dptr(mats), maxdim, -1.0, 1, L.dptr(te), L.dptr(sv),
                                max_keep, L.iptr(kept))
    bpc._psi_dirty = True
    return rc, te, sv, kept


def test_noalloc_apply_matches_the_allocating_entry_point():
    g = itnbp.named_grid((6, 6))
    psi = itnbp.random_tensornetwork(np.random.default_rng(41), np.float64, g, link_space=4)
    a = itnbp.update(itnbp.initialize_cache(psi), maxiter=100, res_tol=1e-12, edge_sequence="parallel")
    b = a.copy()
    deg = {v: g.degree(v) for v in g.vertices}
    classes = edge_coloring(g)
    # first call: a whole colour class; second call: fewer gates of another class
    for edges in (classes[0], classes[1][:5]):
        gates = [O.expm_sym(O.ising_bond_hamiltonian(2.5, deg[u], deg[v]), -0.05) for (u, v) in edges]
        print(f"apply {len(edges)} gates through both entry points", flush=True)
        ra = call(a, "itnbp_apply_gates", edges, gates, 4)
        rb = call(b, "itnbp_apply_gates_noalloc", edges, gates, 4)
        assert ra[0] == L.OK and rb[0] == L.OK, (ra[0], rb[0], a.lib.itnbp_last_error(a.ctx.h))
        for x, y in zip(ra[1:], rb[1:]):
            np.testing.assert_array_equal(x, y)
        assert (ra[1] > 0).all() and (ra[3] == 4).all()
        ta, tb = a.tensornetwork(), b.tensornetwork()
        for v in g.vertices:
            np.testing.assert_array_equal(ta[v], tb[v])


def test_noalloc_apply_refuses_calls_beyond_its_workspace():
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(2), np.float64, g, link_space=2)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=50, res_tol=1e-12, edge_sequence="parallel")
    n = 129
    ids = np.zeros(n, dtype=np.int32)
    mats = np.zeros(16 * n)
    te = np.zeros(n)
    rc = bpc.lib.itnbp_apply_gates_noalloc(bpc.h, n, L.iptr(ids), L.dptr(mats), 2, -1.0, 0, L.dptr(te), None, 0,
                                           None)
    assert rc == L.ERR_UNSUPPORTED
    assert b"at most 128 gates" in bpc.lib.itnbp_last_error(bpc.ctx.h)
    e = (g.vertices[0], g.vertices[1])
    ids = np.array([g.edge_id(*e)], dtype=np.int32)
    sv = np.zeros(33)
    rc = bpc.lib.itnbp_apply_gates_noalloc(bpc.h, 1, L.iptr(ids), L.dptr(np.eye(4).reshape(-1)), 2, -1.0, 0,
                                           L.dptr(te), L.dptr(sv), 33, None)
    assert rc == L.ERR_UNSUPPORTED


def test_noalloc_apply_keeps_bond_dimensions_and_refuses_growth():
    """Without allocation the bonds cannot be resized: a call whose kept rank
    would exceed the bond is refused before any launch, leaving psi unchanged."""
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(6), np.float64, g, link_space=2)
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=50, res_tol=1e-12, edge_sequence="parallel")
    before = {v: np.array(bpc[v]) for v in g.vertices}
    G = np.random.default_rng(3).standard_normal((4, 4))
    rc, te, sv, kept = call(bpc, "itnbp_apply_gates_noalloc", [((1, 1), (2, 1))], [G], 0)
    assert rc == L.ERR_UNSUPPORTED
    for v in g.vertices:
        np.testing.assert_array_equal(bpc[v], before[v])
