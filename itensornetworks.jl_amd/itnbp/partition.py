"""BP over generalised partitions: several vertices per partition.

The reference partitions the quadratic-form network with
``BeliefPropagationCache(tn; partitioned_vertices)`` (src/caches/beliefpropagationcache.jl:20-35,
used by test/test_expect.jl:22-39 to group a grid by columns).  BP on a
partitioning is BP on the quotient graph whose "site tensor" of a partition is
the contraction of its vertices' tensors over the bonds inside it: messages
then live on the bundles of bonds between partitions.  Here that contraction
is done once on the host (``np.einsum`` over the partition's ket tensors), and
the partition network runs through the same device kernels as any network:

* the physical index of a partition combines its members' physical indices,
  first member fastest (column-major);
* the bond between two partitions combines the lattice bonds that cross
  between them, ordered by edge id, first fastest -- the same order on both
  sides;
* a member's one-site RDM is the partial trace of its partition's RDM, also
  taken on the device.

Partition sizes are bounded by the device's generic kernels (combined bond
dimension <= 64, degree <= 6) and by the contraction's limits (2^28 elements
per tensor, 2^30 per call).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .network import ITensorNetwork, NamedGraph

_QF_TAGS = ("ket", "bra", "operator")


def device_einsum(tensors, labels, out_labels, *, ctx=None):
    """Contraction of real tensors on the device (itnbp_einsum): ``labels[i]``
    lists hashable labels of the axes of ``tensors[i]``; a label repeated in
    one tensor is traced, one shared by several and absent from
    ``out_labels`` is contracted, one in a single tensor and not in the
    output is summed.  Returns the array over ``out_labels``."""
    ctx = ctx or L.default_context()
    lib = ctx.lib
    ids, dims = {}, []
    flat, ranks, data = [], [], []
    for t, lab in zip(tensors, labels):
        t = np.asarray(t, dtype=np.float64)
        if t.ndim != len(lab):
            raise ValueError(f"a tensor of rank {t.ndim} has {len(lab)} labels")
        for x, n in zip(lab, t.shape):
            if x not in ids:
                ids[x] = len(ids)
                dims.append(int(n))
            elif dims[ids[x]] != n:
                raise ValueError(f"label {x!r} has dimensions {dims[ids[x]]} and {n}")
            flat.append(ids[x])
        ranks.append(t.ndim)
        data.append(t.reshape(-1, order="F"))
    missing = [x for x in out_labels if x not in ids]
    if missing:
        raise ValueError(f"output labels {missing} appear in no tensor")
    outl = np.array([ids[x] for x in out_labels], dtype=np.int32)
    shape = [dims[ids[x]] for x in out_labels]
    out = np.empty(int(np.prod(shape)) if shape else 1, dtype=np.float64)
    ranks = np.array(ranks, dtype=np.int32)
    flat = np.array(flat if flat else [0], dtype=np.int32)
    dims_a = np.array(dims if dims else [1], dtype=np.int32)
    data = np.concatenate(data) if data else np.zeros(1)
    ctx.check(lib.itnbp_einsum(ctx.h, len(tensors), L.iptr(ranks), L.iptr(flat), L.iptr(dims_a), len(dims),
                               L.dptr(data), len(outl), L.iptr(outl), L.dptr(out)))
    return out.reshape(shape, order="F") if shape else out.reshape(())


def _original_vertex(x, vertices):
    """Map a quadratic-form vertex ``(v, "ket"|"bra"|"operator")`` to v."""
    if x in vertices:
        return x
    if isinstance(x, tuple) and len(x) == 2 and x[1] in _QF_TAGS:
        return x[0]
    raise ValueError(f"{x!r} is not a vertex of the network")


class PartitionedNetwork:
    """``psi`` grouped by ``partitioned_vertices`` (a dict key -> vertices, or a
    list of vertex lists; quadratic-form vertices are accepted and mapped to
    their network vertex).  ``network`` is the partition network."""

    def __init__(self, psi: ITensorNetwork, partitioned_vertices, *, ctx=None):
        g = psi.graph
        self._ctx = ctx
        vset = set(g.vertices)
        groups = partitioned_vertices
        if not isinstance(groups, dict):
            groups = {i: list(vs) for i, vs in enumerate(groups)}
        members, owner = {}, {}
        for key, vs in groups.items():
            lst = []
            for x in vs:
                v = _original_vertex(x, vset)
                if v in owner and owner[v] != key:
                    raise ValueError(f"vertex {v!r} is in two partitions")
                if v not in owner:
                    owner[v] = key
                    lst.append(v)
            if lst:
                members[key] = lst
        missing = [v for v in g.vertices if v not in owner]
        if missing:
            raise ValueError(f"vertices without a partition: {missing[:5]}")
        self.psi = psi
        self.members = members
        self.owner = owner
        keys = list(members)
        self.keys = keys
        # bundles of crossing edges, by partition pair in key order
        bundles = {}
        for e, (a, b) in enumerate(g.edges):
            pa, pb = owner[a], owner[b]
            if pa == pb:
                continue
            k = (pa, pb) if keys.index(pa) < keys.index(pb) else (pb, pa)
            bundles.setdefault(k, []).append(e)
        qedges = sorted(bundles, key=lambda k: bundles[k][0])
        self.bundles = bundles
        qg = NamedGraph(keys, qedges)
        self.graph = qg
        tensors = {key: self._contract(key) for key in keys}
        self.network = ITensorNetwork(qg, tensors)

    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = L.default_context()
        return self._ctx

    def _bundle(self, p, q):
        k = (p, q) if (p, q) in self.bundles else (q, p)
        return self.bundles[k]

    def _contract(self, key):
        psi, g = self.psi, self.psi.graph
        mem = self.members[key]
        tensors, labels = [], []
        for v in mem:
            tensors.append(np.asarray(psi[v]))
            labels.append([("s", v)] + [("e", g.edge_id(v, w)) for w in g.incident(v)])
        out = [("s", v) for v in mem]
        bshape = []
        for q in self.graph.incident(key):
            es = self._bundle(key, q)
            out += [("e", e) for e in es]
            bshape.append(int(np.prod([psi.bond_dim(*g.edges[e]) for e in es])))
        t = device_einsum(tensors, labels, out, ctx=self.ctx)
        # combine: physical indices (first member fastest), then each bundle
        return t.reshape([int(np.prod([psi.phys_dim(v) for v in mem]))] + bshape, order="F")

    def member_rdm(self, rho_p, v):
        """One-site RDM of member ``v`` from its partition's RDM (partial trace)."""
        mem = self.members[self.owner[v]]
        ds = [self.psi.phys_dim(w) for w in mem]
        m = len(mem)
        i = mem.index(v)
        r = np.asarray(rho_p).reshape(ds + ds, order="F")
        # trace out every other member: its ket and bra axes share a label
        lab = [("k", j) for j in range(m)] + [("k", j) if j != i else ("b", j) for j in range(m)]
        return device_einsum([r], [lab], [("k", i), ("b", i)], ctx=self.ctx)
