"""CPU FP64 oracle for BP on two-layer form networks (bra != ket, real or complex).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker, never by the
product (``itnbp``).

The network is <phi|psi> (``LinearFormNetwork``, formnetworks/linearformnetwork.jl:15-25,
built by ``inner_network``, abstractitensornetwork.jl:509-511) with one {bra, ket}
partition per site; <phi|A|psi> (``BilinearFormNetwork``,
formnetworks/bilinearformnetwork.jl:23-42) enters as <phi|A psi> with the
operator's and the ket's link indices of every edge merged.  Conventions as in
``bp_oracle``: site tensors ``t[v].shape == (d, chi_1, ..., chi_k)`` in the
graph's incident order, messages ``M[(u, v)][b_ket, b_bra]``, the bra layer
holds phi itself and is conjugated here.

Pinning (tests/test_oracle.py): on trees one forest-cover sweep makes
``logscalar`` equal the exact log of <phi|psi> and <phi|A|psi> (whole-network
contraction), with phi = psi real every function equals its ``bp_oracle``
counterpart (pinned there), and a complex state that is a phase times a real
state has the real state's messages.
"""
from __future__ import annotations

import numpy as np

from oracle import bp_oracle as O


def open_contraction2(g, ket, bra, msgs, u, open_axes):
    """sum ket_u * conj(bra_u) * prod M_{w->u} over every index not in
    ``open_axes`` (abstractbeliefpropagationcache.jl:225-239 with bra != ket):
    T[(open ket idx), (open bra idx)]."""
    t = ket[u]
    phi = t
    for a, (w, _) in enumerate(g.incident[u], start=1):
        if a in open_axes:
            continue
        phi = O.mode_product(phi, msgs[(w, u)], a)  # ket bond -> bra bond
    b = bra[u]
    rest = [a for a in range(t.ndim) if a not in open_axes]
    T = np.tensordot(phi, np.conj(b), axes=(rest, rest))
    if not open_axes:
        return T.reshape(1, 1)
    ax = open_axes[0]
    return T.reshape(t.shape[ax], b.shape[ax])


def updated_message2(g, ket, bra, msgs, u, v, normalize=True):
    """updated_message on the two-layer network, divided by its Frobenius norm
    unless zero (abstractbeliefpropagationcache.jl:225-239)."""
    T = open_contraction2(g, ket, bra, msgs, u, [g.axis(u, v)])
    n = np.linalg.norm(T)
    return T / n if normalize and n != 0 else T


def bp_update2(g, ket, bra, msgs, maxiter, tol=None, schedule="sequential", sequence=None, normalize=True):
    """update(Algorithm"bp", bpc) (abstractbeliefpropagationcache.jl:272-337) on the
    two-layer network; returns (msgs, iterations, mean message_diff or None)."""
    if maxiter is None:
        raise ValueError("You need to specify a number of iterations for BP!")
    if sequence is None:
        sequence = O.forest_cover_sequence(g) if schedule == "sequential" else O.parallel_sequence(g)
    msgs = dict(msgs)
    it, mean_diff = 0, None
    for it in range(1, maxiter + 1):
        diff = 0.0
        new_msgs = {}
        for (u, v) in sequence:
            new = updated_message2(g, ket, bra, msgs, u, v, normalize)
            if tol is not None:
                diff += O.message_diff(new, msgs[(u, v)])
            if schedule == "sequential":
                msgs[(u, v)] = new
            else:
                new_msgs[(u, v)] = new
        msgs.update(new_msgs)
        if tol is not None:
            mean_diff = diff / len(sequence)
            if mean_diff <= tol:
                break
    return msgs, it, mean_diff


def identity_messages2(g, ket):
    """delta(b_ket, b_bra) on every directed edge (equal bond dimensions)."""
    msgs = {}
    for (u, v) in g.edges:
        chi = ket[u].shape[g.axis(u, v)]
        msgs[(u, v)] = np.eye(chi, dtype=ket[u].dtype)
        msgs[(v, u)] = np.eye(chi, dtype=ket[u].dtype)
    return msgs


def vertex_scalar2(g, ket, bra, msgs, u):
    """region_scalar(bpc, ::QuotientVertex) (beliefpropagationcache.jl:107-113)."""
    return open_contraction2(g, ket, bra, msgs, u, [])[0, 0]


def logscalar2(g, ket, bra, msgs):
    """logscalar(bpc) (abstractbeliefpropagationcache.jl:397-412), complex logarithms."""
    num = [vertex_scalar2(g, ket, bra, msgs, u) for u in range(g.nv)]
    den = [O.edge_scalar(msgs, u, v) for (u, v) in g.edges]
    return complex(np.sum(np.log(np.asarray(num, dtype=complex))) - np.sum(np.log(np.asarray(den, dtype=complex))))


def rdm1_2(g, ket, bra, msgs, u):
    """R[s_ket, s_bra] of site u (environment of beliefpropagationcache.jl:100-105
    contracted with the ket and the conjugated bra), unnormalised."""
    return open_contraction2(g, ket, bra, msgs, u, [0])


def fold_operator(g, psi, A):
    """ket'_v[s', c] = sum_s A_v[s', s, a] psi_v[s, b] with c = b + chi_psi a on
    every bond: <phi|A|psi> as the two-layer network <phi|A psi>."""
    out = []
    for v in range(g.nv):
        p, a = psi[v], A[v]
        k = p.ndim - 1
        # einsum over s; axes (s', b_1, a_1, b_2, a_2, ...)
        letters = "bcdefghijklm"
        pb = letters[:k]
        ab = letters[k:2 * k]
        spec = "Zs" + ab + ",s" + pb + "->Z" + "".join(x + y for x, y in zip(pb, ab))
        t = np.einsum(spec, a, p)
        shape = [t.shape[0]] + [p.shape[1 + i] * a.shape[2 + i] for i in range(k)]
        out.append(t.reshape(shape, order="F"))
    return out


def exact_inner(g, ket, bra):
    """<phi|psi> by contracting both networks into full states."""
    return complex(np.vdot(O.exact_state(g, bra), O.exact_state(g, ket)))


def exact_operator_inner(g, bra, A, ket):
    """<phi|A|psi> by contracting the operator network into a d^N x d^N matrix."""
    cur = np.ones(())
    labels = []
    for v in range(g.nv):
        t = A[v]
        lv = [("o", v), ("i", v)] + [("b", e) for (_, e) in g.incident[v]]
        shared = [lab for lab in lv if lab in labels]
        ia = [labels.index(lab) for lab in shared]
        ib = [lv.index(lab) for lab in shared]
        cur = np.tensordot(cur, t, axes=(ia, ib))
        labels = [lab for lab in labels if lab not in shared] + [lab for lab in lv if lab not in shared]
    perm = [labels.index(("o", v)) for v in range(g.nv)] + [labels.index(("i", v)) for v in range(g.nv)]
    Aop = np.transpose(cur, perm)
    dims = Aop.shape[:g.nv]
    N = int(np.prod(dims))
    Amat = Aop.reshape(N, N)
    k = O.exact_state(g, ket).reshape(N)
    b = O.exact_state(g, bra).reshape(N)
    return complex(np.vdot(b, Amat @ k))
