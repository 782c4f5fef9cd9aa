"""Generalised partitions (BeliefPropagationCache(tn; partitioned_vertices),
test/test_expect.jl:22-39): the device contraction of each partition gives
the exact state back; column-grouped BP on the device is exact on grids (the
partition graph is a path) and singleton partitions reproduce plain BP; the
device contraction itself (itnbp_einsum) against numpy's einsum."""
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


@pytest.mark.gpu
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


def _einsum_cases():
    rng = np.random.default_rng(21)
    r = lambda *s: rng.standard_normal(s)
    # (operands as (array, labels) pairs, output labels)
    yield "matrix product", [(r(5, 7), "ab"), (r(7, 3), "bc")], "ac"
    yield "chain with a batch label", [(r(4, 6, 5), "abz"), (r(6, 3, 5), "bcz"), (r(3, 2), "cd")], "adz"
    yield "hyperedge", [(r(4, 5), "ab"), (r(5, 6), "bc"), (r(5, 2), "bd")], "acd"
    yield "trace and sum", [(r(3, 4, 3, 2), "abac"), (r(4, 5), "bd")], "d"
    yield "outer product", [(r(3, 2), "ab"), (r(4,), "c")], "cab"
    yield "scalar", [(r(6, 5), "ab"), (r(6, 5), "ab")], ""
    yield "permutation only", [(r(2, 3, 4, 5), "abcd")], "dbca"
    yield "partial trace", [(r(2, 3, 2, 3), "abAb")], "aA"
    yield "grid patch", [(r(2, 3, 4), "sab"), (r(2, 3, 5, 4), "tacd"), (r(2, 4, 6), "ubx"), (r(2, 5, 6, 3), "vcxy")], \
        "stuvdy"
    yield "large product", [(r(70, 130), "ab"), (r(130, 90), "bc")], "ca"


@pytest.mark.gpu
@pytest.mark.parametrize("name,ops,out", list(_einsum_cases()), ids=[c[0] for c in _einsum_cases()])
def test_device_einsum_matches_numpy(name, ops, out):
    from itnbp.partition import device_einsum
    got = device_einsum([a for a, _ in ops], [list(lab) for _, lab in ops], list(out))
    want = np.einsum(",".join(lab for _, lab in ops) + "->" + out, *[a for a, _ in ops])
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.gpu
def test_device_einsum_errors():
    from itnbp.partition import device_einsum
    with pytest.raises(ValueError):
        device_einsum([np.ones((2, 3))], [["a", "b"]], ["c"])      # output label in no tensor
    with pytest.raises(itnbp.ITNBPError):
        device_einsum([np.ones((2, 3))], [["a", "b"]], ["a", "a"])  # repeated output label
