"""CPU FP64 oracle for the belief-propagation (BP) hot path of ITensorNetworks.jl.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and by nothing else.  The product (``itnbp``) never calls it and
fails loudly when its HIP library is missing.

It restates, in plain NumPy, the reference algorithm on the <psi|psi> norm
network (a QuadraticFormNetwork with one {operator, bra, ket} partition per
lattice site).  Every function cites the reference file:line it follows
(paths relative to the ITensorNetworks.jl repository root).

Conventions (shared with the C ABI):
  * a network is ``Graph(nv, edges)``; the bonds of site tensor ``psi[v]`` are
    ordered like ``incident[v]`` (neighbours by ascending vertex index unless
    an explicit order is given), after the physical axis:
    ``psi[v].shape == (d, chi_{e1}, ..., chi_{ek})``
    (``test/utils.jl:32-36``: ``inds_v = [siteinds[v]; link_v]``);
  * a message ``M[(u, v)]`` is a ``chi x chi`` array ``M[b_ket, b_bra]``
    (``quadraticformnetwork.jl:115-119``: ``delta(dag(map_idx(k)), k)``);
  * one-site operators are ``O[s', s]`` (ITensors ``op(name, s)``); two-site
    operators are ``d^2 x d^2`` matrices ``O[(s1', s2'), (s1, s2)]`` with the
    first site's index fastest.

Pinning (what checks each function, beyond restating it): see the module
``tests/test_oracle.py`` - exact contraction on small lattices, tree
exactness of one forest-cover sweep, product states, closed-form gates,
LAPACK SVD/eigh identities, and brute-force einsum of the full message.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np

EPS64 = float(np.finfo(np.float64).eps)


# ----------------------------------------------------------------------------
# graphs
# ----------------------------------------------------------------------------
class Graph:
    """Undirected simple graph with a fixed bond (axis) order per vertex.

    ``incident[v]`` is a list of ``(w, e)`` pairs; position ``a`` in that list
    is tensor axis ``a + 1`` of ``psi[v]`` (axis 0 is physical).
    """

    def __init__(self, nv, edges, incident=None):
        self.nv = int(nv)
        self.edges = [tuple(int(x) for x in e) for e in edges]
        eid = {}
        for i, (u, v) in enumerate(self.edges):
            if u == v:
                raise ValueError("self loop")
            key = (min(u, v), max(u, v))
            if key in eid:
                raise ValueError("multi-edge")  # itensornetwork.jl:150-153 (hyperedge ban)
            eid[key] = i
        self._eid = eid
        if incident is None:
            incident = [[] for _ in range(self.nv)]
            for i, (u, v) in enumerate(self.edges):
                incident[u].append((v, i))
                incident[v].append((u, i))
            incident = [sorted(lst) for lst in incident]
        self.incident = [list(map(tuple, lst)) for lst in incident]

    def edge_id(self, u, v):
        return self._eid[(min(u, v), max(u, v))]

    def axis(self, v, w):
        """Tensor axis (1-based after the physical axis) of psi[v] facing w."""
        for a, (x, _) in enumerate(self.incident[v]):
            if x == w:
                return a + 1
        raise KeyError((v, w))

    def neighbors(self, v):
        return [w for (w, _) in self.incident[v]]

    def directed_edges(self):
        """``vcat(edges(g), reverse.(edges(g)))`` (edge_sequences.jl:49-51)."""
        return list(self.edges) + [(v, u) for (u, v) in self.edges]

    def is_tree(self):
        if len(self.edges) != self.nv - 1:
            return False
        return len(_components(self.nv, self.edges)) == 1


def _components(nv, edges):
    adj = [[] for _ in range(nv)]
    for u, v in edges:
        adj[u].append(v)
        adj[v].append(u)
    seen, comps = [False] * nv, []
    for r in range(nv):
        if seen[r]:
            continue
        comp, q = [], deque([r])
        seen[r] = True
        while q:
            x = q.popleft()
            comp.append(x)
            for y in sorted(adj[x]):
                if not seen[y]:
                    seen[y] = True
                    q.append(y)
        comps.append(sorted(comp))
    return comps


def square_lattice(Lx, Ly, periodic=False):
    """Open (or periodic) Lx x Ly grid; vertex (i, j) -> i + Lx*j, i fastest
    (``named_grid((Lx, Ly))`` column-major vertex order); edges (u, v), u < v,
    sorted."""
    idx = lambda i, j: i + Lx * j
    es = set()
    for j in range(Ly):
        for i in range(Lx):
            for (ii, jj) in ((i + 1, j), (i, j + 1)):
                if periodic:
                    ii, jj = ii % Lx, jj % Ly
                if ii < Lx and jj < Ly and (ii, jj) != (i, j):
                    a, b = idx(i, j), idx(ii, jj)
                    es.add((min(a, b), max(a, b)))
    return Graph(Lx * Ly, sorted(es))


def heavy_hex_lattice(Lx, Ly):
    """Heavy-hex lattice: a brick-wall honeycomb on an Lx x Ly grid (all
    horizontal bonds, vertical bonds where i + j is even) with one extra
    vertex inserted on every bond."""
    idx = lambda i, j: i + Lx * j
    base = []
    for j in range(Ly):
        for i in range(Lx):
            if i + 1 < Lx:
                base.append((idx(i, j), idx(i + 1, j)))
            if j + 1 < Ly and (i + j) % 2 == 0:
                base.append((idx(i, j), idx(i, j + 1)))
    nv = Lx * Ly
    es = []
    for (u, v) in sorted(base):
        m = nv
        nv += 1
        es.append((u, m))
        es.append((v, m))
    return Graph(nv, sorted((min(a, b), max(a, b)) for a, b in es))


def path_graph(n):
    return Graph(n, [(i, i + 1) for i in range(n - 1)])


def comb_tree(nx, ny):
    """Spine along the first row, teeth along columns."""
    idx = lambda i, j: i + nx * j
    es = [(idx(i, 0), idx(i + 1, 0)) for i in range(nx - 1)]
    es += [(idx(i, j), idx(i, j + 1)) for i in range(nx) for j in range(ny - 1)]
    return Graph(nx * ny, sorted((min(a, b), max(a, b)) for a, b in es))


# ----------------------------------------------------------------------------
# schedules
# ----------------------------------------------------------------------------
def forest_cover_sequence(g):
    """Sequential BP edge schedule (edge_sequences.jl:32-47).

    The edges are covered by spanning forests (the first a BFS spanning forest
    of the whole graph, the next one of the edges left over, and so on; NamedGraphs'
    ``forest_cover`` is not vendored, so this cover is our own, deterministic
    one).  For every tree of every forest, rooted at its lowest vertex, the
    post-order DFS edges (child -> parent) are followed by their reverses in
    reverse order (``vcat(tree_edges, reverse(reverse.(tree_edges)))``).
    Every directed edge appears exactly once per sweep.
    """
    remaining = list(g.edges)
    seq = []
    while remaining:
        adj = [[] for _ in range(g.nv)]
        for (u, v) in remaining:
            adj[u].append(v)
            adj[v].append(u)
        for lst in adj:
            lst.sort()
        seen = [False] * g.nv
        forest = set()
        children = [[] for _ in range(g.nv)]
        roots = []
        for r in range(g.nv):
            if seen[r] or not adj[r]:
                continue
            seen[r] = True
            roots.append(r)
            q = deque([r])
            while q:
                x = q.popleft()
                for y in adj[x]:
                    if not seen[y]:
                        seen[y] = True
                        children[x].append(y)
                        forest.add((min(x, y), max(x, y)))
                        q.append(y)
        for r in roots:
            post = []

            def dfs(x, parent):
                for c in children[x]:
                    dfs(c, x)
                if parent is not None:
                    post.append((x, parent))

            dfs(r, None)
            seq += post + [(b, a) for (a, b) in reversed(post)]
        remaining = [(u, v) for (u, v) in remaining if (min(u, v), max(u, v)) not in forest]
    return seq


def parallel_sequence(g):
    """``[[e] for e in vcat(edges(g), reverse.(edges(g)))]`` (edge_sequences.jl:49-51)."""
    return g.directed_edges()


# ----------------------------------------------------------------------------
# tensors
# ----------------------------------------------------------------------------
def random_peps(g, d, chi, rng, dtype=np.float64, scale=1.0):
    """iid N(0, scale^2) site tensors, shape (d, chi_e...) in incident order
    (test/utils.jl:23-38 with our own seeded generator).  ``chi`` is an int
    or a per-edge list."""
    chis = [chi] * len(g.edges) if np.isscalar(chi) else list(chi)
    psi = []
    for v in range(g.nv):
        shape = (d,) + tuple(chis[e] for (_, e) in g.incident[v])
        psi.append((rng.standard_normal(shape) * scale).astype(dtype))
    return psi


def mode_product(X, M, ax):
    """Y[..., c, ...] = sum_x X[..., x, ...] * M[x, c] on axis ``ax``."""
    Y = np.tensordot(X, M, axes=([ax], [0]))
    return np.moveaxis(Y, -1, ax)


def identity_messages(g, psi):
    """delta(b_bra, b_ket) on every directed edge (quadraticformnetwork.jl:96-124)."""
    msgs = {}
    for (u, v) in g.edges:
        chi = psi[u].shape[g.axis(u, v)]
        msgs[(u, v)] = np.eye(chi, dtype=psi[u].dtype)
        msgs[(v, u)] = np.eye(chi, dtype=psi[u].dtype)
    return msgs


def initialize_messages(g, psi):
    """initialize_cache(Algorithm"bp", ::QuadraticFormNetwork) (initialize_cache.jl:14-29):
    empty on trees, identity messages on loopy graphs."""
    return {} if g.is_tree() else identity_messages(g, psi)


def _absorb(g, psi_u, msgs, u, skip_axes):
    phi = psi_u
    for a, (w, _) in enumerate(g.incident[u], start=1):
        if a in skip_axes:
            continue
        phi = mode_product(phi, msgs[(w, u)], a)
    return phi


def open_contraction(g, psi, msgs, u, open_axes):
    """T[(open ket idx), (open bra idx)] = sum psi_u * conj(psi_u) * prod M_{w->u}
    over every bond axis not in ``open_axes`` (axis 0 = physical, always summed
    unless open).  The message update (open = {axis to v}), the one-site RDM
    (open = {0}) and the vertex scalar (open = {}) are all this contraction
    (abstractbeliefpropagationcache.jl:225-239, beliefpropagationcache.jl:100-113)."""
    t = psi[u]
    phi = _absorb(g, t, msgs, u, set(open_axes))
    rest = [a for a in range(t.ndim) if a not in open_axes]
    T = np.tensordot(phi, np.conj(t), axes=(rest, rest))
    n_open = len(open_axes)
    # tensordot leaves (phi open axes..., bra open axes...) in increasing axis order
    rows = int(np.prod([t.shape[a] for a in sorted(open_axes)])) if n_open else 1
    return T.reshape((rows, rows), order="F") if n_open else T.reshape(1, 1)


def updated_message(g, psi, msgs, u, v, normalize=True):
    """updated_message(Algorithm"contract", bpc, e) (abstractbeliefpropagationcache.jl:225-239):
    contract the incoming messages except the reverse edge with the ket and bra
    tensors of ``u``, then divide by the Frobenius norm unless it is zero."""
    T = open_contraction(g, psi, msgs, u, [g.axis(u, v)])
    n = np.linalg.norm(T)
    if normalize and n != 0:
        T = T / n
    return T


def message_diff(a, b):
    """1 - |<a/|a|, b/|b|>|^2 (abstractbeliefpropagationcache.jl:32-36)."""
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    f = abs(np.vdot(a / na, b / nb)) ** 2
    return 1.0 - f


def rel_residual(new, old):
    """||new - old||_F / ||new||_F: the residual that certifies parity below the
    ~1e-8 floor of ``message_diff`` (which is quadratic in the angle)."""
    n = np.linalg.norm(new)
    return float(np.linalg.norm(new - old) / n) if n else 0.0


def bp_update(g, psi, msgs, maxiter, tol=None, schedule="sequential", sequence=None,
              normalize=True, verbose=False):
    """update(Algorithm"bp", bpc) (abstractbeliefpropagationcache.jl:313-329) with the
    sequential sweep (:272-287) or the intended Jacobi sweep (:294-308: every group
    computed from the previous cache, then merged).

    Returns (msgs, iterations run, mean message_diff of the last sweep or None,
    max rel residual of the last sweep)."""
    if maxiter is None:
        raise ValueError("You need to specify a number of iterations for BP!")  # :315-317
    if sequence is None:
        sequence = forest_cover_sequence(g) if schedule == "sequential" else parallel_sequence(g)
    msgs = dict(msgs)
    it, mean_diff, max_res = 0, None, 0.0
    for it in range(1, maxiter + 1):
        diff = 0.0
        max_res = 0.0
        if schedule == "sequential":
            for (u, v) in sequence:
                new = updated_message(g, psi, msgs, u, v, normalize)
                if tol is not None:
                    old = msgs[(u, v)]  # message() has no default (:173-175)
                    diff += message_diff(new, old)
                    max_res = max(max_res, rel_residual(new, old))
                msgs[(u, v)] = new
        else:
            new_msgs = {}
            for (u, v) in sequence:
                new = updated_message(g, psi, msgs, u, v, normalize)
                if tol is not None:
                    old = msgs[(u, v)]
                    diff += message_diff(new, old)
                    max_res = max(max_res, rel_residual(new, old))
                new_msgs[(u, v)] = new
            msgs.update(new_msgs)
        if tol is not None:
            mean_diff = diff / len(sequence)
            if mean_diff <= tol:  # :321
                if verbose:
                    print(f"BP converged to desired precision after {it} iterations.")
                break
    return msgs, it, mean_diff, max_res


def bp_update_groups(g, psi, msgs, groups, maxiter, tol=None, normalize=True):
    """update over edge groups (abstractbeliefpropagationcache.jl:294-308, 313-329):
    in every sweep each group is updated sequentially (:272-287) starting from the
    messages the sweep started from, the groups' messages are merged, and the
    summed message_diff is divided by the number of groups (``length`` of the
    sequence, :321).  Returns (msgs, iterations, mean diff)."""
    if maxiter is None:
        raise ValueError("You need to specify a number of iterations for BP!")
    msgs = dict(msgs)
    it, mean = 0, None
    for it in range(1, maxiter + 1):
        merged, diff = {}, 0.0
        for grp in groups:
            work = dict(msgs)
            for (u, v) in grp:
                new = updated_message(g, psi, work, u, v, normalize)
                diff += message_diff(new, work[(u, v)])
                work[(u, v)] = new
            for e in grp:
                merged[e] = work[e]
        msgs.update(merged)
        mean = diff / len(groups)
        if tol is not None and mean <= tol:
            break
    return msgs, it, mean


# ----------------------------------------------------------------------------
# environments, expectation values, scalars
# ----------------------------------------------------------------------------
def rdm1(g, psi, msgs, u):
    """rho_u[s, s'] = sum psi_u[s, b] conj(psi_u[s', b']) prod_w M_{w->u}[b_w, b'_w]
    (environment(bpc, [operator vertex]) contracted with bra and ket, expect.jl:10-16)."""
    return open_contraction(g, psi, msgs, u, [0])


def expect1(op, rho):
    """numerator/denominator of expect.jl:5-19: sum O[s', s] rho[s, s'] / sum rho[s, s]."""
    return np.sum(op.T * rho) / np.trace(rho)


def half_pair(g, psi, msgs, u, v):
    """A_u[(s, b), (s', b')] with the physical axis and the bond to v open."""
    return open_contraction(g, psi, msgs, u, [0, g.axis(u, v)])


def rdm2(g, psi, msgs, u, v):
    """Two-site RDM of the pair (u, v) from the BP environment of all three layers
    of both sites (test/test_belief_propagation.jl:64-91):
    rho[(s_u, s_v), (s_u', s_v')], first site fastest."""
    d_u, d_v = psi[u].shape[0], psi[v].shape[0]
    chi = psi[u].shape[g.axis(u, v)]
    Au = half_pair(g, psi, msgs, u, v).reshape(d_u, chi, d_u, chi, order="F")
    Av = half_pair(g, psi, msgs, v, u).reshape(d_v, chi, d_v, chi, order="F")
    rho = np.einsum("ibjc,kblc->ikjl", Au, Av)
    return rho.reshape(d_u * d_v, d_u * d_v, order="F")


def expect2(op, rho):
    return np.sum(op.T * rho) / np.trace(rho)


def vertex_scalar(g, psi, msgs, u):
    """region_scalar(bpc, ::QuotientVertex) (beliefpropagationcache.jl:107-113)."""
    return open_contraction(g, psi, msgs, u, [])[0, 0]


def edge_scalar(msgs, u, v):
    """region_scalar(bpc, ::QuotientEdge) (beliefpropagationcache.jl:115-119)."""
    return np.sum(msgs[(u, v)] * msgs[(v, u)])


def logscalar(g, psi, msgs):
    """logscalar(bpc) (abstractbeliefpropagationcache.jl:397-408)."""
    num = np.array([vertex_scalar(g, psi, msgs, u) for u in range(g.nv)])
    den = np.array([edge_scalar(msgs, u, v) for (u, v) in g.edges])
    if np.any(np.real(num) < 0):
        num = num.astype(complex)
    if np.any(np.real(den) < 0):
        den = den.astype(complex)
    if np.any(den == 0):
        return -np.inf
    with np.errstate(divide="ignore"):
        return np.sum(np.log(num)) - np.sum(np.log(den))


def rescale_messages(g, msgs):
    """rescale_messages(bpc, pes) (beliefpropagationcache.jl:121-139)."""
    msgs = dict(msgs)
    for (u, v) in g.edges:
        me = msgs[(u, v)] / np.linalg.norm(msgs[(u, v)])
        mer = msgs[(v, u)] / np.linalg.norm(msgs[(v, u)])
        n = np.sum(me * mer)
        if np.isrealobj(n) or np.imag(n) == 0:
            n = np.real(n)
            me = me * np.sign(n)
            n = n * np.sign(n)
        sf = 1.0 / math.sqrt(n)
        msgs[(u, v)] = sf * me
        msgs[(v, u)] = sf * mer
    return msgs


def rescale_partitions(g, psi, msgs):
    """rescale_partitions(bpc, partitions; verts = ket and bra vertices)
    (abstractbeliefpropagationcache.jl:349-379) as used by normalize(alg="bp")
    (normalize.jl:125-142): every ket (and bra) tensor is divided by its norm, then
    both by Z_v^{-1/2} so that every vertex scalar becomes 1."""
    out = []
    for u in range(g.nv):
        t = psi[u] / np.linalg.norm(psi[u])
        tmp = list(psi)
        tmp[u] = t
        zv = np.real(vertex_scalar(g, tmp, msgs, u))
        s = np.sign(zv)
        out.append(t * (s * abs(zv) ** (-0.5)) if s < 0 else t * zv ** (-0.5))
    return out


def normalize_bp(g, psi, msgs):
    """normalize(alg="bp") with a converged cache (normalize.jl:125-142):
    rescale_messages then rescale_partitions; returns (psi', msgs')."""
    msgs = rescale_messages(g, msgs)
    return rescale_partitions(g, psi, msgs), msgs


# ----------------------------------------------------------------------------
# exact contraction (oracle role only; contract.jl:7-21)
# ----------------------------------------------------------------------------
def exact_state(g, psi):
    """Contract the whole ket network into a d^N state tensor (site order 0..N-1)."""
    cur = np.ones(())
    labels = []
    for v in range(g.nv):
        t = psi[v]
        lv = [("s", v)] + [("b", e) for (_, e) in g.incident[v]]
        shared = [l for l in lv if l in labels]
        ia = [labels.index(l) for l in shared]
        ib = [lv.index(l) for l in shared]
        cur = np.tensordot(cur, t, axes=(ia, ib))
        labels = [l for l in labels if l not in shared] + [l for l in lv if l not in shared]
    perm = [labels.index(("s", v)) for v in range(g.nv)]
    return np.transpose(cur, perm)


def exact_rdm1(state, v):
    X = np.moveaxis(state, v, 0).reshape(state.shape[v], -1)
    return X @ X.conj().T  # rho[s, s']


def exact_rdm2(state, u, v):
    X = np.moveaxis(state, [u, v], [0, 1])
    du, dv = X.shape[0], X.shape[1]
    X = X.reshape(du, dv, -1)
    rho = np.einsum("abr,cdr->abcd", X, X.conj())  # [su, sv, su', sv']
    return rho.reshape(du * dv, du * dv, order="F")  # rows (su, sv), su fastest


def exact_norm2(state):
    return float(np.vdot(state, state).real)


# ----------------------------------------------------------------------------
# operators
# ----------------------------------------------------------------------------
SZ = np.array([[0.5, 0.0], [0.0, -0.5]])
X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.array([[1.0, 0.0], [0.0, -1.0]])
I2 = np.eye(2)


def two_site(A, B):
    """O[(s1', s2'), (s1, s2)] with the first site fastest = kron(B, A)."""
    return np.kron(B, A)


def ising_bond_hamiltonian(h, deg_u, deg_v, J=1.0):
    """H_b = J Z(x)Z + h (X(x)I / deg_u + I(x)X / deg_v): the transverse field split
    over the bonds of each site so that sum_b H_b = H."""
    return J * two_site(Z, Z) + h * (two_site(X, I2) / deg_u + two_site(I2, X) / deg_v)


def expm_sym(H, t):
    """exp(t H) for a real symmetric H."""
    w, U = np.linalg.eigh(H)
    return (U * np.exp(t * w)) @ U.T


# ----------------------------------------------------------------------------
# BP-gauged two-site gate application (apply.jl:9-146)
# ----------------------------------------------------------------------------
def truncation_keep(p, cutoff, maxdim=None):
    """Number of (descending) weights ``p`` to keep: drop from the small end while
    the dropped weight / total stays <= cutoff, then cap at maxdim; keep >= 1.
    Returns (keep, dropped fraction)."""
    p = np.asarray(p, dtype=float)
    tot = float(np.sum(p))
    n = len(p)
    keep = n
    if maxdim is not None:
        keep = min(keep, int(maxdim))
    if cutoff is not None and tot > 0:
        disc = float(np.sum(p[keep:]))
        while keep > 1 and (disc + p[keep - 1]) / tot <= cutoff:
            disc += p[keep - 1]
            keep -= 1
    keep = max(keep, 1)
    err = float(np.sum(p[keep:]) / tot) if tot > 0 else 0.0
    return keep, err


def map_eigvals(f, A, cutoff=10 * EPS64):
    """map_eigvals(f, A, L, R; ishermitian=true, cutoff) (apply.jl:9-25): Hermitian
    eigendecomposition A = U D U^dag, eigenvalues truncated by ``cutoff`` on
    |lambda| (sorted by magnitude), then U f(D) U^dag."""
    A = 0.5 * (A + A.conj().T)
    w, U = np.linalg.eigh(A)
    order = np.argsort(-np.abs(w), kind="stable")
    w, U = w[order], U[:, order]
    keep, _ = truncation_keep(np.abs(w), cutoff)
    w, U = w[:keep], U[:, :keep]
    return (U * f(w)) @ U.conj().T


def simple_update_bp(G, psi_u, psi_v, ax_u, ax_v, envs_u, envs_v, maxdim=None,
                     cutoff=None, normalize=False):
    """simple_update_bp (apply.jl:33-95) + the normalize step of apply (:134-137).

    ``envs_u`` maps every bond axis of psi_u except ``ax_u`` to its incoming message
    M_{w->u}; likewise ``envs_v``.  ``G`` is the d^2 x d^2 gate O[(su', sv'), (su, sv)].
    Returns (psi_u', psi_v', singular values kept, truncation error); the bond
    between u and v keeps its axis position, with the new dimension.
    """
    du, dv = psi_u.shape[0], psi_v.shape[0]

    def reduce_side(t, ax, envs):
        S = {a: map_eigvals(np.sqrt, M) for a, M in envs.items()}
        Sinv = {a: map_eigvals(lambda x: 1.0 / np.sqrt(x), M) for a, M in envs.items()}
        tt = t
        for a, s in S.items():
            tt = mode_product(tt, s, a)
        ext = [a for a in range(1, t.ndim) if a != ax]
        tp = np.transpose(tt, ext + [0, ax])
        E = int(np.prod([t.shape[a] for a in ext])) if ext else 1
        Xm = tp.reshape(E, t.shape[0] * t.shape[ax], order="F")
        Q, R = np.linalg.qr(Xm, mode="reduced")  # apply.jl:74-75
        r = Q.shape[1]
        R = R.reshape(r, t.shape[0], t.shape[ax], order="F")
        return Q, R, ext, Sinv

    Qu, Ru, ext_u, Sinv_u = reduce_side(psi_u, ax_u, envs_u)
    Qv, Rv, ext_v, Sinv_v = reduce_side(psi_v, ax_v, envs_v)
    ru, rv = Ru.shape[0], Rv.shape[0]
    # Theta[ru, su, rv, sv] = sum G[(su, sv), (su0, sv0)] Ru[ru, su0, b] Rv[rv, sv0, b]  (apply.jl:78)
    G4 = G.reshape(du, dv, du, dv, order="F")  # [su', sv', su, sv]
    RR = np.einsum("aib,cjb->aicj", Ru, Rv)
    Th = np.einsum("klij,aicj->akcl", G4, RR)
    Tm = Th.reshape(ru * du, rv * dv, order="F")
    U, s, Vh = np.linalg.svd(Tm, full_matrices=False)  # factorize_svd(ortho="none"), :81-88
    keep, truncerr = truncation_keep(s ** 2, cutoff, maxdim)
    U, s, Vh = U[:, :keep], s[:keep], Vh[:keep]
    sq = np.sqrt(s)
    Ru2 = (U * sq).reshape(ru, du, keep, order="F")
    Rv2 = (sq[:, None] * Vh).T.reshape(rv, dv, keep, order="F")

    def rebuild(t, ax, Q, R2, ext, Sinv):
        # psi'[ext..., s, b_new] = sum_r Q[ext, r] R2[r, s, b_new]; then Sinv on each ext bond
        core = (Q @ R2.reshape(R2.shape[0], -1, order="F"))
        shape = [t.shape[a] for a in ext] + [t.shape[0], R2.shape[2]]
        core = core.reshape(shape, order="F")
        # back to [s, bonds in axis order] with the uv bond at position ax
        nd = t.ndim
        src_axes = ext + [0, ax]
        out = np.moveaxis(core, list(range(nd)), src_axes)
        for a, si in Sinv.items():
            out = mode_product(out, si.T, a)  # sum_{b'} Q[..b'..] Sinv[b, b']
        return out

    pu = rebuild(psi_u, ax_u, Qu, Ru2, ext_u, Sinv_u)
    pv = rebuild(psi_v, ax_v, Qv, Rv2, ext_v, Sinv_v)
    # deterministic gauge of the new bond: the largest-magnitude entry of every
    # slice of psi_u' (first in column-major order) is made positive, and the
    # same sign is applied to psi_v' (the pair product is unchanged)
    for j in range(keep):
        sl = np.take(pu, j, axis=ax_u).reshape(-1, order="F")
        if np.sign(sl[np.argmax(np.abs(sl))]) < 0:
            idx_u = [slice(None)] * pu.ndim
            idx_u[ax_u] = j
            pu[tuple(idx_u)] *= -1
            idx_v = [slice(None)] * pv.ndim
            idx_v[ax_v] = j
            pv[tuple(idx_v)] *= -1
    if normalize:
        pu = pu / np.linalg.norm(pu)
        pv = pv / np.linalg.norm(pv)
    return pu, pv, s, truncerr


def envs_for_bond(g, msgs, u, v):
    """The product environment of the pair (u, v): the messages into u from every
    neighbour but v, keyed by axis of psi_u, and likewise for v
    (environment(bpc, all layers of [u, v]), beliefpropagationcache.jl:100-105)."""
    eu = {g.axis(u, w): msgs[(w, u)] for w in g.neighbors(u) if w != v}
    ev = {g.axis(v, w): msgs[(w, v)] for w in g.neighbors(v) if w != u}
    return eu, ev


def apply_gate(G, g, psi, msgs, u, v, maxdim=None, cutoff=None, normalize=False):
    """ITensors.apply(o, psi; envs, maxdim, cutoff, normalize) for a two-site gate
    (apply.jl:97-146).  Returns (new psi list, singular values, truncerr)."""
    if g.edge_id(u, v) is None:
        raise ValueError("Vertices where the gates are being applied must be neighbors for now.")
    eu, ev = envs_for_bond(g, msgs, u, v)
    pu, pv, s, err = simple_update_bp(G, psi[u], psi[v], g.axis(u, v), g.axis(v, u),
                                      eu, ev, maxdim, cutoff, normalize)
    out = list(psi)
    out[u], out[v] = pu, pv
    return out, s, err


def edge_coloring(g):
    """Greedy proper edge colouring: classes of disjoint bonds."""
    colors = []
    for (u, v) in g.edges:
        for c in colors:
            if all(u not in (a, b) and v not in (a, b) for (a, b) in c):
                c.append((u, v))
                break
        else:
            colors.append([(u, v)])
    return colors


def simple_update_round(g, psi, msgs, gate_fn, maxdim, cutoff=None, normalize=True):
    """One round of BP-gauged gates over every bond, colour class by colour class,
    every gate with envs from the same (converged) cache."""
    errs, svals = {}, {}
    for cls in edge_coloring(g):
        for (u, v) in cls:
            psi, s, err = apply_gate(gate_fn(u, v), g, psi, msgs, u, v, maxdim, cutoff, normalize)
            errs[(u, v)], svals[(u, v)] = err, s
    return psi, errs, svals


def bp_ising_energy(g, psi, msgs, h, J=1.0):
    """E = J sum_<uv> <Z_u Z_v> + h sum_v <X_v> from BP RDMs."""
    e = 0.0
    for (u, v) in g.edges:
        e += J * expect2(two_site(Z, Z), rdm2(g, psi, msgs, u, v))
    for u in range(g.nv):
        e += h * expect1(X, rdm1(g, psi, msgs, u))
    return float(np.real(e))
