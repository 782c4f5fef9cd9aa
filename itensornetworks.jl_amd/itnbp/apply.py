"""BP-gauged gate application: ITensors.apply(o, psi; envs, maxdim, cutoff,
normalize, callback) of src/apply.jl:97-160, on the device.

A gate is ``Gate(matrix, sites)``: one site (d x d, O[s', s]) or two
neighbouring sites ((d_u d_v) x (d_u d_v), O[(s_u', s_v'), (s_u, s_v)], s_u
fastest).  Two-site gates need a product environment: a BeliefPropagationCache
(its messages into the pair) or the dict returned by ``environment(bpc, [u, v])``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .cache import BeliefPropagationCache, QuadraticFormNetwork, update_
from .network import ITensorNetwork


@dataclass
class Gate:
    matrix: np.ndarray
    sites: tuple


def _reorder_two_site(m, du, dv):
    """O[(su', sv'), (su, sv)] -> O[(sv', su'), (sv, su)]."""
    return m.reshape(du, dv, du, dv, order="F").transpose(1, 0, 3, 2).reshape(du * dv, du * dv, order="F")


def apply_gates_(bpc: BeliefPropagationCache, gates, *, maxdim=None, cutoff=None, normalize=False,
                 callback=None):
    """Apply one- and two-site gates in place on the cache's site tensors (one-site
    gates on the device, apply.jl:110-116), every two-site gate with
    the cache's current messages as its (product) environment.  Gates are grouped
    greedily, in order, into launches of disjoint bonds.  Returns the list of
    truncation errors; ``callback(singular_values=..., truncation_error=...)`` is
    called per gate (apply.jl:89)."""
    g = bpc.graph
    # physical dimensions only (gates keep them): the host description, with no
    # download of the site tensors the previous gates changed
    ket = bpc.qf.ket
    pending = list(enumerate(gates))
    errs = [None] * len(gates)
    while pending:
        used, batch, site_batch, rest = set(), [], [], []
        for i, gt in pending:
            sites = tuple(gt.sites)
            if len(sites) == 1:
                if sites[0] in used:
                    rest.append((i, gt))
                else:
                    used.add(sites[0])
                    site_batch.append((i, gt))
                continue
            if len(sites) != 2:
                raise ValueError("Gates with more than 2 sites is not supported yet.")
            u, v = sites
            if not g.has_edge(u, v):
                raise ValueError("Vertices where the gates are being applied must be neighbors for now.")
            if u in used or v in used:
                rest.append((i, gt))
            else:
                used.update((u, v))
                batch.append((i, gt))
        pending = rest
        if site_batch:
            _apply_site_gates(bpc, site_batch, normalize, errs)
        if not batch:
            continue
        ids, mats = [], []
        max_keep = 64  # the largest kept rank (bonds grow to at most 64)
        for i, gt in batch:
            u, v = gt.sites
            e = g.edge_id(u, v)
            m = np.asarray(gt.matrix, dtype=np.float64)
            du, dv = ket.phys_dim(u), ket.phys_dim(v)
            if m.shape != (du * dv, du * dv):
                raise ValueError("gate shape does not match the physical dimensions")
            if g.edges[e] != (u, v):
                m = _reorder_two_site(m, du, dv)
            ids.append(e)
            mats.append(m.reshape(-1, order="F"))
        ids = np.array(ids, dtype=np.int32)
        mats = np.ascontiguousarray(np.concatenate(mats))
        te = np.empty(len(batch), dtype=np.float64)
        sv = np.zeros(len(batch) * max_keep, dtype=np.float64)
        kept = np.empty(len(batch), dtype=np.int32)
        try:
            bpc.ctx.check(bpc.lib.itnbp_apply_gates(
                bpc.h, len(batch), L.iptr(ids), L.dptr(mats), int(maxdim or 0),
                -1.0 if cutoff is None else float(cutoff), int(bool(normalize)), L.dptr(te), L.dptr(sv),
                max_keep, L.iptr(kept)))
        finally:
            # a failed call may still have changed the tensors of other bonds
            bpc._psi_dirty = True
        for k, (i, gt) in enumerate(batch):
            errs[i] = float(te[k])
            if callback is not None:
                callback(singular_values=sv[k * max_keep:k * max_keep + kept[k]].copy(),
                         truncation_error=float(te[k]))
    return errs


def _apply_site_gates(bpc, site_batch, normalize, errs):
    """One-site gates on distinct vertices, on the device (apply.jl:110-116); their
    truncation error is 0."""
    g = bpc.graph
    verts = np.array([g.index[gt.sites[0]] for _, gt in site_batch], dtype=np.int32)
    mats = []
    for _, gt in site_batch:
        d = bpc.qf.ket.phys_dim(gt.sites[0])
        m = np.asarray(gt.matrix, dtype=np.float64)
        if m.shape != (d, d):
            raise ValueError("gate shape does not match the physical dimension")
        mats.append(m.reshape(-1, order="F"))
    mats = np.ascontiguousarray(np.concatenate(mats))
    try:
        bpc.ctx.check(bpc.lib.itnbp_apply_site_gates(bpc.h, len(verts), L.iptr(verts), L.dptr(mats),
                                                     int(bool(normalize)), None))
    finally:
        bpc._psi_dirty = True
    for i, _ in site_batch:
        errs[i] = 0.0


def _check_product_env(envs, u, v):
    for (w, x) in envs:
        if x not in (u, v) or w in (u, v):
            raise RuntimeError(
                "`apply` requires a product environment (`envs` with no shared edges); "
                "contract `envs` to product form before calling.")


def apply(o: Gate, psi: ITensorNetwork, *, envs=None, maxdim=None, cutoff=None, normalize=False,
          callback=None):
    """apply(o, psi; envs, maxdim, cutoff, normalize, callback) (apply.jl:97-146):
    returns a new ITensorNetwork."""
    sites = tuple(o.sites)
    if len(sites) == 1:
        ctx = envs.ctx if isinstance(envs, BeliefPropagationCache) else None
        bpc = BeliefPropagationCache(QuadraticFormNetwork(psi), storage="f64", ctx=ctx)
        apply_gates_(bpc, [o], normalize=normalize)
        return bpc.tensornetwork()
    if len(sites) > 2:
        raise ValueError("Gates with more than 2 sites is not supported yet.")
    if len(sites) < 1:
        raise ValueError("Gate being applied does not share indices with tensor network.")
    u, v = sites
    if not psi.graph.has_edge(u, v):
        raise ValueError("Vertices where the gates are being applied must be neighbors for now.")
    if isinstance(envs, BeliefPropagationCache):
        bpc = BeliefPropagationCache(QuadraticFormNetwork(psi), storage="f64", ctx=envs.ctx)
        msgs = {e: m for e, m in envs.messages().items() if e[1] in (u, v) and e[0] not in (u, v)}
    else:
        msgs = dict(envs or {})
        _check_product_env(msgs, u, v)
        bpc = BeliefPropagationCache(QuadraticFormNetwork(psi), storage="f64")
    bpc.set_messages_(msgs)
    apply_gates_(bpc, [o], maxdim=maxdim, cutoff=cutoff, normalize=normalize, callback=callback)
    return bpc.tensornetwork()


def edge_coloring(graph):
    """Greedy proper edge colouring in edge order: classes of disjoint bonds.
    Computed once per graph (graphs are not modified after construction) and
    kept on it."""
    cached = getattr(graph, "_edge_classes", None)
    if cached is None:
        cached, used = [], []
        for (u, v) in graph.edges:
            for c, us in zip(cached, used):
                if u not in us and v not in us:
                    c.append((u, v))
                    us.update((u, v))
                    break
            else:
                cached.append([(u, v)])
                used.append({u, v})
        graph._edge_classes = cached
    return [list(c) for c in cached]


def simple_update_round_(bpc: BeliefPropagationCache, gate_fn, *, maxdim, cutoff=None, normalize=True):
    """One BP-gauged gate on every bond, colour class by colour class, all with the
    cache's current messages as environments.  Returns {edge: truncerr}."""
    errs = {}
    for cls in edge_coloring(bpc.graph):
        gates = [Gate(gate_fn(u, v), (u, v)) for (u, v) in cls]
        te = apply_gates_(bpc, gates, maxdim=maxdim, cutoff=cutoff, normalize=normalize)
        errs.update(zip(cls, te))
    return errs


def tebd_bp_(bpc: BeliefPropagationCache, gate_fn, nsteps, *, maxdim, cutoff=None, normalize=True,
             bp_maxiter=30, bp_tol=1e-10, edge_sequence="parallel", callback=None):
    """Gate evolution with BP regauging between rounds, in place on the cache.

    The reference's TEBD loop (src/tebd.jl:2-31) applies gates without
    environments; this is the BP-gauged variant: every step first updates the
    cache's messages (warm-started from the previous step's) and then applies
    one gate per bond, colour class by colour class, with those messages as the
    product environment of every gate.  ``gate_fn(u, v)`` returns the
    (d_u d_v)^2 gate of bond (u, v) (first site fastest), e.g. exp(-dt H_uv).
    A final update leaves the cache consistent with the evolved state.
    ``callback(step, truncation_errors)`` is called after every gate round.
    Returns the per-step maximum truncation error."""
    hist = []
    for step in range(int(nsteps)):
        update_(bpc, maxiter=bp_maxiter, tol=bp_tol, edge_sequence=edge_sequence)
        errs = simple_update_round_(bpc, gate_fn, maxdim=maxdim, cutoff=cutoff, normalize=normalize)
        hist.append(max(errs.values()) if errs else 0.0)
        if callback is not None:
            callback(step, errs)
    update_(bpc, maxiter=bp_maxiter, tol=bp_tol, edge_sequence=edge_sequence)
    return hist


def apply_gates_vjp(bpc: BeliefPropagationCache, gates, kept, dpsi_out, *, maxdim=None, cutoff=None,
                    normalize=False):
    """Vector-Jacobian product of ``apply_gates_`` for two-site gates on disjoint
    bonds (simple_update_bp, apply.jl:33-95, and the normalisation of :134-137).

    Call it on the cache as it was before the gates (site tensors and messages
    unchanged), with the arguments of the forward call, the ranks it kept and the
    cotangents ``dpsi_out[i] = (dpsi_u', dpsi_v')`` of the updated tensors of gate
    i (``gates[i].sites = (u, v)``; the new bond has dimension ``kept[i]``).
    Returns ``(dpsi, dgates, dmsgs)``: per gate the cotangents of psi_u and psi_v
    and of the gate matrix, and for every environment message used, its
    cotangent (a dict by directed edge)."""
    g = bpc.graph
    ket = bpc.tensornetwork()
    ids, mats, outs, flips = [], [], [], []
    used = set()
    for gt, (cu, cv) in zip(gates, dpsi_out):
        u, v = gt.sites
        if u in used or v in used:
            raise ValueError("the bonds of one apply_gates_vjp call must be disjoint")
        used.update((u, v))
        e = g.edge_id(u, v)
        m = np.asarray(gt.matrix, dtype=np.float64)
        du, dv = ket.phys_dim(u), ket.phys_dim(v)
        flip = g.edges[e] != (u, v)
        if flip:
            m = _reorder_two_site(m, du, dv)
            cu, cv = cv, cu
        ids.append(e)
        mats.append(m.reshape(-1, order="F"))
        outs += [np.asarray(cu, dtype=np.float64).reshape(-1, order="F"),
                 np.asarray(cv, dtype=np.float64).reshape(-1, order="F")]
        flips.append(flip)
    n = len(ids)
    ids = np.array(ids, dtype=np.int32)
    mats = np.ascontiguousarray(np.concatenate(mats))
    outs = np.ascontiguousarray(np.concatenate(outs))
    kept = np.ascontiguousarray(kept, dtype=np.int32)
    in_sizes = []
    for e in ids:
        a, b = g.edges[e]
        in_sizes += [ket[a].size, ket[b].size]
    din = np.zeros(sum(in_sizes))
    dg = np.zeros(mats.size)
    dm = np.zeros(int(bpc.lib.itnbp_net_msg_numel(bpc.h)))
    bpc.ctx.check(bpc.lib.itnbp_apply_vjp(
        bpc.h, n, L.iptr(ids), L.dptr(mats), int(maxdim or 0), -1.0 if cutoff is None else float(cutoff),
        int(normalize), L.iptr(kept), L.dptr(outs), L.dptr(din), L.dptr(dg), L.dptr(dm)))
    dpsi, dgates, off, goff = [], [], 0, 0
    for i, e in enumerate(ids):
        a, b = g.edges[e]
        da = din[off:off + ket[a].size].reshape(ket[a].shape, order="F")
        off += ket[a].size
        db = din[off:off + ket[b].size].reshape(ket[b].shape, order="F")
        off += ket[b].size
        dd = ket.phys_dim(a) * ket.phys_dim(b)
        gm = dg[goff:goff + dd * dd].reshape(dd, dd, order="F")
        goff += dd * dd
        if flips[i]:
            gm = _reorder_two_site(gm, ket.phys_dim(a), ket.phys_dim(b))
            da, db = db, da
        dpsi.append((da, db))
        dgates.append(gm)
    dmsgs = {}
    used_msgs = set()
    for e in ids:
        a, b = g.edges[e]
        for x, y in ((a, b), (b, a)):
            for w in g.incident(x):
                if w != y:
                    used_msgs.add((w, x))
    off = 0
    for mid in range(2 * g.ne):
        a, b = g.message_edge(mid)
        chi = ket.bond_dim(a, b)
        if (a, b) in used_msgs:
            dmsgs[(a, b)] = dm[off:off + chi * chi].reshape(chi, chi, order="F")
        off += chi * chi
    return dpsi, dgates, dmsgs
