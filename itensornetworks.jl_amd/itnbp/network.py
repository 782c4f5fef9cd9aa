"""Host-side network model: named graphs, lattices and ITensorNetwork-like
containers of dense site tensors.

A site tensor of vertex ``v`` is a NumPy array of shape
``(d, chi_1, ..., chi_k)`` whose bond axes follow ``graph.incident(v)``
(neighbours by ascending vertex index unless given explicitly), like the
reference's ``inds_v = [siteinds[v]; link_v]`` (test/utils.jl:32-36).
"""
from __future__ import annotations

from collections import deque

import numpy as np


class NamedGraph:
    """Undirected simple graph over hashable vertex names (insertion ordered)."""

    def __init__(self, vertices, edges, incident=None):
        self.vertices = list(vertices)
        self.index = {v: i for i, v in enumerate(self.vertices)}
        if len(self.index) != len(self.vertices):
            raise ValueError("duplicate vertex")
        es = []
        seen = set()
        for (a, b) in edges:
            ia, ib = self.index[a], self.index[b]
            if ia == ib:
                raise ValueError("self loop")
            key = (min(ia, ib), max(ia, ib))
            if key in seen:
                raise ValueError("multiple edges between the same vertices are not supported")
            seen.add(key)
            es.append((a, b))
        self.edges = es
        self._eid = {}
        for i, (a, b) in enumerate(es):
            self._eid[(a, b)] = i
            self._eid[(b, a)] = i
        if incident is None:
            inc = {v: [] for v in self.vertices}
            for (a, b) in es:
                inc[a].append(b)
                inc[b].append(a)
            incident = {v: sorted(lst, key=lambda w: self.index[w]) for v, lst in inc.items()}
        self._inc = {v: list(incident[v]) for v in self.vertices}

    @property
    def nv(self):
        return len(self.vertices)

    @property
    def ne(self):
        return len(self.edges)

    def incident(self, v):
        """Neighbours of v in bond-axis order."""
        return self._inc[v]

    def neighbors(self, v):
        return list(self._inc[v])

    def degree(self, v):
        return len(self._inc[v])

    def edge_id(self, a, b):
        return self._eid[(a, b)]

    def has_edge(self, a, b):
        return (a, b) in self._eid

    def axis(self, v, w):
        """1-based tensor axis of psi_v facing w (axis 0 is physical)."""
        return self._inc[v].index(w) + 1

    def message_id(self, a, b):
        """Directed message id of the C ABI: 2e for edges[e] = (a, b), 2e+1 reversed."""
        e = self._eid[(a, b)]
        return 2 * e if self.edges[e] == (a, b) else 2 * e + 1

    def message_edge(self, m):
        a, b = self.edges[m // 2]
        return (a, b) if m % 2 == 0 else (b, a)

    def is_tree(self):
        if self.ne != self.nv - 1:
            return False
        adj = {v: [] for v in self.vertices}
        for a, b in self.edges:
            adj[a].append(b)
            adj[b].append(a)
        seen = {self.vertices[0]}
        q = deque([self.vertices[0]])
        while q:
            x = q.popleft()
            for y in adj[x]:
                if y not in seen:
                    seen.add(y)
                    q.append(y)
        return len(seen) == self.nv

    def csr(self):
        """(edges[2ne], inc_ptr[nv+1], inc_edge[2ne]) integer arrays of the C ABI."""
        edges = np.array([[self.index[a], self.index[b]] for a, b in self.edges],
                         dtype=np.int32).reshape(-1)
        inc_ptr = [0]
        inc_edge = []
        for v in self.vertices:
            for w in self._inc[v]:
                inc_edge.append(self._eid[(v, w)])
            inc_ptr.append(len(inc_edge))
        return edges, np.array(inc_ptr, dtype=np.int32), np.array(inc_edge, dtype=np.int32)

    def directed_edges(self):
        """vcat(edges(g), reverse.(edges(g))) (edge_sequences.jl:49-51)."""
        return list(self.edges) + [(b, a) for (a, b) in self.edges]


def named_grid(dims, periodic=False):
    """``named_grid((Lx, Ly))``: vertices (i, j), 1-based, i fastest; edges sorted by
    their vertex positions."""
    Lx, Ly = dims
    verts = [(i, j) for j in range(1, Ly + 1) for i in range(1, Lx + 1)]
    pos = {v: k for k, v in enumerate(verts)}
    es = set()
    for (i, j) in verts:
        for (ii, jj) in ((i + 1, j), (i, j + 1)):
            if periodic:
                ii = (ii - 1) % Lx + 1
                jj = (jj - 1) % Ly + 1
            if ii <= Lx and jj <= Ly and (ii, jj) != (i, j):
                a, b = (i, j), (ii, jj)
                if pos[a] > pos[b]:
                    a, b = b, a
                es.add((a, b))
    edges = sorted(es, key=lambda e: (pos[e[0]], pos[e[1]]))
    return NamedGraph(verts, edges)


def heavy_hex(Lx, Ly):
    """Brick-wall honeycomb on an Lx x Ly grid with a vertex inserted on every bond.
    Vertices: ("q", i, j) for the lattice sites, ("b", k) for the bond vertices."""
    base = []
    for j in range(Ly):
        for i in range(Lx):
            if i + 1 < Lx:
                base.append(((i, j), (i + 1, j)))
            if j + 1 < Ly and (i + j) % 2 == 0:
                base.append(((i, j), (i, j + 1)))
    verts = [("q", i, j) for j in range(Ly) for i in range(Lx)]
    edges = []
    for k, (a, b) in enumerate(sorted(base, key=lambda e: (e[0][0] + Lx * e[0][1], e[1][0] + Lx * e[1][1]))):
        m = ("b", k)
        verts.append(m)
        edges.append((("q",) + a, m))
        edges.append((("q",) + b, m))
    pos = {v: i for i, v in enumerate(verts)}
    edges = [(a, b) if pos[a] < pos[b] else (b, a) for a, b in edges]
    edges.sort(key=lambda e: (pos[e[0]], pos[e[1]]))
    return NamedGraph(verts, edges)


def path_graph(n):
    return NamedGraph(list(range(1, n + 1)), [(i, i + 1) for i in range(1, n)])


def comb_tree(nx, ny):
    verts = [(i, j) for j in range(1, ny + 1) for i in range(1, nx + 1)]
    es = [((i, 1), (i + 1, 1)) for i in range(1, nx)]
    es += [((i, j), (i, j + 1)) for i in range(1, nx + 1) for j in range(1, ny)]
    pos = {v: k for k, v in enumerate(verts)}
    es = sorted(((a, b) if pos[a] < pos[b] else (b, a) for a, b in es), key=lambda e: (pos[e[0]], pos[e[1]]))
    return NamedGraph(verts, es)


class ITensorNetwork:
    """A tensor network state: a graph plus one dense tensor per vertex."""

    def __init__(self, graph: NamedGraph, tensors):
        self.graph = graph
        self.tensors = {v: np.asarray(tensors[v]) for v in graph.vertices}
        for v in graph.vertices:
            t = self.tensors[v]
            if t.ndim != 1 + graph.degree(v):
                raise ValueError(f"tensor of {v} has rank {t.ndim}, expected {1 + graph.degree(v)}")
        for (a, b) in graph.edges:
            if self.tensors[a].shape[graph.axis(a, b)] != self.tensors[b].shape[graph.axis(b, a)]:
                raise ValueError(f"bond dimension mismatch on edge {(a, b)}")

    def __getitem__(self, v):
        return self.tensors[v]

    def __setitem__(self, v, t):
        t = np.asarray(t)
        if t.shape[:1] != self.tensors[v].shape[:1] or t.ndim != self.tensors[v].ndim:
            raise ValueError("shape change of a site tensor")
        self.tensors[v] = t

    @property
    def vertices(self):
        return self.graph.vertices

    @property
    def dtype(self):
        return np.result_type(*[t.dtype for t in self.tensors.values()])

    def copy(self):
        return ITensorNetwork(self.graph, {v: t.copy() for v, t in self.tensors.items()})

    def bond_dim(self, a, b):
        return self.tensors[a].shape[self.graph.axis(a, b)]

    def phys_dim(self, v):
        return self.tensors[v].shape[0]

    def is_tree(self):
        return self.graph.is_tree()


def random_tensornetwork(rng, eltype, graph, *, link_space=1, phys_dim=2, scale=1.0):
    """Seeded iid Gaussian site tensors (test/utils.jl:23-38 with NumPy's PCG64):
    ``psi_v = scale * randn(d, chi...)`` in vertex order."""
    link = link_space if not np.isscalar(link_space) else [link_space] * graph.ne
    tensors = {}
    for v in graph.vertices:
        shape = (phys_dim,) + tuple(link[graph.edge_id(v, w)] for w in graph.incident(v))
        tensors[v] = (rng.standard_normal(shape) * scale).astype(eltype)
    return ITensorNetwork(graph, tensors)


def random_instance_set(rng, n=256, eltype=np.float32):
    """``n`` independent networks for batched BP: each an L x L open grid (L uniform
    in 6..12) or a heavy-hex lattice (3..6 hexagons a side), d = 2, one bond
    dimension per instance uniform in {4, 6, 8, 12, 16}, entries N(0, 1/chi).
    Returns a list of (kind, chi, ITensorNetwork)."""
    out = []
    for _ in range(n):
        if rng.random() < 0.5:
            L = int(rng.integers(6, 13))
            g, kind = named_grid((L, L)), "grid"
        else:
            L = int(rng.integers(3, 7))
            g, kind = heavy_hex(L, L), "heavy_hex"
        chi = int(rng.choice([4, 6, 8, 12, 16]))
        out.append((kind, chi, random_tensornetwork(rng, eltype, g, link_space=chi, scale=1 / np.sqrt(chi))))
    return out


def productstate(graph, states, eltype=np.float64, phys_dim=2):
    """Product state with bond dimension 1; ``states[v]`` is a basis index or a vector."""
    tensors = {}
    for v in graph.vertices:
        s = states(v) if callable(states) else states[v]
        vec = np.zeros(phys_dim, dtype=eltype)
        if np.isscalar(s):
            vec[int(s)] = 1
        else:
            vec[:] = s
        tensors[v] = vec.reshape((phys_dim,) + (1,) * graph.degree(v))
    return ITensorNetwork(graph, tensors)


def merge_link_indices(graph, tensors, link_dims):
    """Site tensors whose edges carry several link indices, as the one-bond
    network the library runs on.

    ``link_dims[(u, w)]`` lists the dimensions of the edge's link indices in the
    order they appear on both of its tensors; ``tensors[v]`` has the axes
    ``(s, links of its first incident edge..., links of the next...)`` in the
    graph's incident order.  The links of every edge merge into one bond of
    dimension prod(link_dims) (first link fastest), so a message of the
    returned network is the product-index form of the reference's per-edge
    message vector (identity_messages pushes one delta per link index,
    formnetworks/quadraticformnetwork.jl:115-119, whose product is the delta
    of the merged bond); ``split_message`` undoes the merge."""
    dims_of = {}
    for (u, w), dims in link_dims.items():
        dims_of[(u, w)] = tuple(int(x) for x in dims)
        dims_of[(w, u)] = tuple(int(x) for x in dims)
    out = {}
    for v in graph.vertices:
        t = np.asarray(tensors[v])
        shape = [t.shape[0]]
        a = 1
        for w in graph.incident(v):
            dims = dims_of[(v, w)]
            if tuple(t.shape[a:a + len(dims)]) != dims:
                raise ValueError(f"tensor of {v}: link axes toward {w} have shape {t.shape[a:a + len(dims)]}, "
                                 f"expected {dims}")
            shape.append(int(np.prod(dims)))
            a += len(dims)
        if a != t.ndim:
            raise ValueError(f"tensor of {v} has {t.ndim - a} axes beyond its links")
        out[v] = t.reshape(shape, order="F")
    return ITensorNetwork(graph, out)


def split_message(M, dims):
    """A message of a merged bond as the tensor [k_1 .. k_n, k_1' .. k_n'] over
    the edge's link indices (ket links, then bra links)."""
    dims = tuple(int(x) for x in dims)
    return np.asarray(M).reshape(dims + dims, order="F")
