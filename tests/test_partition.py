"""Generalised partitions (BeliefPropagationCache(tn; partitioned_vertices),
test/test_expect.jl:22-39): the host contraction of each partition is checked
against the exact state on the CPU; column-grouped BP on the device is exact
on grids (the partition graph is a path) and singleton partitions reproduce
plain BP."""
import numpy as np
import pytest

import itnbp
from itnbp.partition import PartitionedNetwork
from helpers import oracle_graph, oracle_psi
from oracle import bp_oracle as O


def _columns(g):
    cols = {}
    for v in g.vertices:
        cols.setdefault(v[0], []).append(v)
    return cols


def test_partition_network_has_the_same_state():
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(4), np.float64, g, link_space=2)
    qf_groups = {k: [(v, t) for v in vs for t in ("ket", "bra", "operator")] for k, vs in _columns(g).items()}
    pn = PartitionedNetwork(psi, qf_groups)
    assert [pn.graph.degree(p) for p in pn.keys] == [1, 2, 1]
    full = O.exact_state(oracle_graph(g), oracle_psi(psi))
    part = O.exact_state(oracle_graph(pn.graph), oracle_psi(pn.network))
    # split every partition's physical index into its members (first fastest)
    split = part.reshape([2] * 9, order="F")
    order = [g.index[v] for p in pn.keys for v in pn.members[p]]
    back = np.moveaxis(split, list(range(9)), order)
    np.testing.assert_allclose(back, full, rtol=1e-12, atol=1e-12)


def test_partition_errors():
    g = itnbp.named_grid((2, 2))
    psi = itnbp.random_tensornetwork(np.random.default_rng(0), np.float64, g, link_space=2)
    with pytest.raises(ValueError):
        PartitionedNetwork(psi, {0: [(1, 1), (1, 2)], 1: [(2, 1)]})          # (2, 2) missing
    with pytest.raises(ValueError):
        PartitionedNetwork(psi, {0: [(1, 1), (1, 2)], 1: [(1, 2), (2, 1), (2, 2)]})  # (1, 2) twice


@pytest.mark.gpu
@pytest.mark.parametrize("L", [2, 3])
def test_column_grouped_bp_is_exact_on_grids(L):
    g = itnbp.named_grid((L, L))
    psi = itnbp.random_tensornetwork(np.random.default_rng(1234), np.float64, g, link_space=2)
    qf = [(v, t) for v in g.vertices for t in ("ket", "bra", "operator")]
    groups = {}
    for x in qf:
        groups.setdefault(x[0][0], []).append(x)
    sz = itnbp.expect(psi, "Sz", cache_construction_kwargs=dict(partitioned_vertices=groups),
                      cache_update_kwargs=dict(maxiter=20))
    og, opsi = oracle_graph(g), oracle_psi(psi)
    state = O.exact_state(og, opsi)
    for v in g.vertices:
        exact = O.expect1(O.SZ, O.exact_rdm1(state, g.index[v]))
        assert sz[v] == pytest.approx(exact, abs=1e-10)


@pytest.mark.gpu
def test_singleton_partitions_reproduce_plain_bp():
    g = itnbp.named_grid((3, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(9), np.float64, g, link_space=2)
    plain = itnbp.expect(psi, "Sz", cache_update_kwargs=dict(maxiter=40, edge_sequence="parallel"))
    single = itnbp.expect(psi, "Sz", cache_construction_kwargs=dict(partitioned_vertices=[[v] for v in g.vertices]),
                          cache_update_kwargs=dict(maxiter=40, edge_sequence="parallel"))
    for v in g.vertices:
        assert single[v] == pytest.approx(plain[v], abs=1e-12)
