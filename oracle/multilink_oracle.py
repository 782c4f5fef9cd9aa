"""CPU FP64 oracle for BP on norm networks whose edges carry several link
indices.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker, never by the
product (``itnbp``).

A site tensor ``psi[v]`` has shape ``(d, links of its first incident edge...,
links of the next...)``: every edge e = (u, w) carries ``len(link_dims[e])``
link indices, in the same order at both ends.  A message is a tensor
``M[(u, w)]`` of shape ``link_dims[e] + link_dims[e]`` (ket links, then bra
links); identity_messages puts one delta per link index
(formnetworks/quadraticformnetwork.jl:115-119) and the message update,
message_diff and scalars act on these tensors index by index
(abstractbeliefpropagationcache.jl:32-36, 225-239, 397-412), with no merging of
indices anywhere, so that it checks the host's merge of an edge's links into
one bond.

Pinning (tests/test_layer_oracle.py): on trees one forest-cover sweep gives the
exact <psi|psi> of the whole network; with one link per edge it equals
``bp_oracle``.
"""
from __future__ import annotations

import numpy as np

from oracle import bp_oracle as O


def link_axes(g, link_dims, v):
    """{neighbour: [tensor axes of the edge's links]} of site v (axis 0 physical)."""
    out, a = {}, 1
    for (w, e) in g.incident[v]:
        n = len(link_dims[e])
        out[w] = list(range(a, a + n))
        a += n
    return out


def identity_messages(g, link_dims, dtype=np.float64):
    msgs = {}
    for e, (u, w) in enumerate(g.edges):
        dims = link_dims[e]
        M = np.ones(())
        for k in dims:
            M = np.multiply.outer(M, np.eye(k, dtype=dtype))
        # axes (k1, k1', k2, k2', ...) -> (k1, k2, ..., k1', k2', ...)
        n = len(dims)
        M = np.transpose(M, list(range(0, 2 * n, 2)) + list(range(1, 2 * n, 2)))
        msgs[(u, w)] = M.copy()
        msgs[(w, u)] = M.copy()
    return msgs


def open_contraction(g, link_dims, psi, msgs, u, open_w=None, open_phys=False):
    t = psi[u]
    axes = link_axes(g, link_dims, u)
    X = t
    # absorb: for every other neighbour w, contract X's ket link axes of edge
    # (w, u) with the message's ket axes; the message's bra axes take their place
    for (w, e) in g.incident[u]:
        if w == open_w:
            continue
        n = len(link_dims[e])
        M = msgs[(w, u)]
        ax = axes[w]
        X = np.tensordot(X, M, axes=(ax, list(range(n))))
        # the n new axes are last: move them back to the link positions
        X = np.moveaxis(X, list(range(X.ndim - n, X.ndim)), ax)
    keep = ([0] if open_phys else []) + (axes[open_w] if open_w is not None else [])
    rest = [a for a in range(t.ndim) if a not in keep]
    return np.tensordot(X, np.conj(t), axes=(rest, rest))


def updated_message(g, link_dims, psi, msgs, u, w, normalize=True):
    T = open_contraction(g, link_dims, psi, msgs, u, open_w=w)
    n = np.linalg.norm(T)
    return T / n if normalize and n != 0 else T


def bp_update(g, link_dims, psi, msgs, maxiter, schedule="sequential"):
    seq = O.forest_cover_sequence(g) if schedule == "sequential" else O.parallel_sequence(g)
    msgs = dict(msgs)
    for _ in range(maxiter):
        new_msgs = {}
        for (u, w) in seq:
            new = updated_message(g, link_dims, psi, msgs, u, w)
            if schedule == "sequential":
                msgs[(u, w)] = new
            else:
                new_msgs[(u, w)] = new
        msgs.update(new_msgs)
    return msgs


def logscalar(g, link_dims, psi, msgs):
    num = [open_contraction(g, link_dims, psi, msgs, u).reshape(()) for u in range(g.nv)]
    den = [np.sum(msgs[(u, w)] * msgs[(w, u)]) for (u, w) in g.edges]
    return complex(np.sum(np.log(np.asarray(num, dtype=complex))) - np.sum(np.log(np.asarray(den, dtype=complex))))


def exact_norm2(g, link_dims, psi):
    """<psi|psi> by contracting the whole ket network (each link its own label)."""
    cur = np.ones(())
    labels = []
    for v in range(g.nv):
        t = psi[v]
        lv = [("s", v)]
        for (_, e) in g.incident[v]:
            lv += [("l", e, i) for i in range(len(link_dims[e]))]
        shared = [lab for lab in lv if lab in labels]
        ia = [labels.index(lab) for lab in shared]
        ib = [lv.index(lab) for lab in shared]
        cur = np.tensordot(cur, t, axes=(ia, ib))
        labels = [lab for lab in labels if lab not in shared] + [lab for lab in lv if lab not in shared]
    return float(np.vdot(cur, cur).real)
