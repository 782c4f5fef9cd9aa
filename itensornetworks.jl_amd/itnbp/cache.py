"""BeliefPropagationCache and the BP API of ITensorNetworks.jl, on the MI355X
library.

Mirrors (paths relative to the ITensorNetworks.jl repository):
  BeliefPropagationCache / copy / messages     src/caches/beliefpropagationcache.jl:13-47
  initialize_cache                             src/initialize_cache.jl:10-29
  identity_messages                            src/formnetworks/quadraticformnetwork.jl:96-124
  update / updated_message / message_diff      src/caches/abstractbeliefpropagationcache.jl:32-36, 225-337
  environment                                  src/caches/beliefpropagationcache.jl:100-105, src/environment.jl:6-49
  region_scalar / logscalar / scalar           src/caches/beliefpropagationcache.jl:107-119,
                                               src/caches/abstractbeliefpropagationcache.jl:397-412
  rescale / normalize                          src/caches/beliefpropagationcache.jl:121-139,
                                               src/caches/abstractbeliefpropagationcache.jl:339-395,
                                               src/normalize.jl:125-142
  expect                                       src/expect.jl:5-109

The functional API returns new caches (a device-side copy); ``update_`` and the
``*_`` setters mutate in place, like the reference's ``!`` variants.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .network import ITensorNetwork

# ITensors op(name, s) for S=1/2 sites, O[s', s]
OPS = {
    "Sz": np.array([[0.5, 0.0], [0.0, -0.5]]),
    "Sx": np.array([[0.0, 0.5], [0.5, 0.0]]),
    "X": np.array([[0.0, 1.0], [1.0, 0.0]]),
    "Z": np.array([[1.0, 0.0], [0.0, -1.0]]),
    "N": np.array([[0.0, 0.0], [0.0, 1.0]]),
    "Id": np.eye(2),
    "I": np.eye(2),
}

PRECISIONS = {"f64": L.PREC_F64, "f32": L.PREC_F32, "bf16_f64": L.PREC_BF16_F64, "bf16x3": L.PREC_BF16X3,
              "bf16x3_f64": L.PREC_BF16X3_F64, "bf16_bf16x3": L.PREC_BF16_BF16X3,
              "bf16_bf16x3_f64": L.PREC_BF16_BF16X3_F64, "f16x2": L.PREC_F16X2, "bf16_f16x2": L.PREC_BF16_F16X2}


def op_matrix(op, d=2):
    if isinstance(op, str):
        if op not in OPS:
            raise ValueError(f"operator {op!r} is not available (real S=1/2 operators only)")
        m = OPS[op]
    else:
        m = np.asarray(op, dtype=np.float64)
    if m.shape != (d, d):
        raise ValueError(f"operator shape {m.shape} does not match the physical dimension {d}")
    return m


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even); returns (float32 values, uint16 bits)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    bits = r.astype(np.uint16)
    back = (bits.astype(np.uint32) << 16).view(np.float32)
    return back, bits


class QuadraticFormNetwork:
    """<psi|I|psi> (src/formnetworks/quadraticformnetwork.jl:75-89): the ket is
    ``psi``, the bra is conj(psi) with every index primed, the operator layer
    is the identity.  Only the ket is stored; the device builds the bra on the
    fly and elides the identity."""

    def __init__(self, ket: ITensorNetwork):
        self.ket = ket

    def ket_network(self):
        return self.ket

    @property
    def graph(self):
        return self.ket.graph

    def is_tree(self):
        return self.ket.graph.is_tree()


def norm_sqr_network(psi):
    return QuadraticFormNetwork(psi)


def _flat_sites(psi: ITensorNetwork, storage):
    parts = [np.asarray(psi[v]).reshape(-1, order="F") for v in psi.graph.vertices]
    flat = np.concatenate(parts) if parts else np.zeros(0)
    if storage == L.F64:
        return np.ascontiguousarray(flat, dtype=np.float64)
    if storage == L.F32:
        return np.ascontiguousarray(flat, dtype=np.float32)
    return bf16_round(flat)[1]


def _unflat_sites(psi_like: ITensorNetwork, flat, storage, shapes=None):
    if storage == L.BF16:
        flat = (flat.astype(np.uint32) << 16).view(np.float32)
    out = {}
    off = 0
    for v in psi_like.graph.vertices:
        shape = psi_like[v].shape if shapes is None else shapes[v]
        n = int(np.prod(shape))
        out[v] = np.array(flat[off:off + n]).reshape(shape, order="F")
        off += n
    return ITensorNetwork(psi_like.graph, out)


class BeliefPropagationCache:
    """BP cache of the norm network of ``psi`` held on one GPU.

    ``storage``: "f64", "f32" or "bf16" (site tensors rounded to bfloat16);
    defaults to the tensors' dtype.
    """

    def __init__(self, tn, messages=None, *, storage=None, ctx=None, _handle=None):
        qf = tn if isinstance(tn, QuadraticFormNetwork) else QuadraticFormNetwork(tn)
        self.qf = qf
        self.ctx = ctx or L.default_context()
        self.lib = self.ctx.lib
        psi = qf.ket
        g = psi.graph
        if storage is None:
            storage = "f32" if psi.dtype == np.float32 else "f64"
        if np.iscomplexobj(np.zeros(0, dtype=psi.dtype)):
            raise ValueError("complex site tensors run on the two-layer engine: build the cache with "
                             "initialize_cache(psi) (an itnbp.LayeredBPCache)")
        self.storage = {"f64": L.F64, "f32": L.F32, "bf16": L.BF16}[storage]
        self.storage_name = storage
        self._host_psi = psi
        self._psi_dirty = False
        self.last_update = None
        self.partitioning = None  # PartitionedNetwork when built over generalised partitions
        if _handle is not None:
            self.h = _handle
            self.ctx._nets.add(self)
            return
        edges, inc_ptr, inc_edge = g.csr()
        chis = np.array([psi.bond_dim(a, b) for (a, b) in g.edges], dtype=np.int32)
        phys = np.array([psi.phys_dim(v) for v in g.vertices], dtype=np.int32)
        h = C.c_void_p()
        self.ctx.check(self.lib.itnbp_net_create(
            self.ctx.h, g.nv, g.ne, L.iptr(edges) if g.ne else None, L.iptr(inc_ptr),
            L.iptr(inc_edge) if g.ne else None, L.iptr(chis) if g.ne else None, L.iptr(phys),
            self.storage, C.byref(h)))
        self.h = h
        self.ctx._nets.add(self)
        flat = _flat_sites(psi, self.storage)
        self.ctx.check(self.lib.itnbp_set_sites(self.h, L.vptr(flat)))
        if self.storage == L.BF16:
            self._host_psi = _unflat_sites(psi, flat, L.BF16)
        if messages:
            self.set_messages_(messages)

    # -- lifetime ------------------------------------------------------------
    def _release(self):
        if getattr(self, "h", None):
            self.lib.itnbp_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def copy(self):
        """Deep copy (src/caches/beliefpropagationcache.jl:43-47)."""
        h = C.c_void_p()
        self.ctx.check(self.lib.itnbp_net_copy(self.h, C.byref(h)))
        c = BeliefPropagationCache.__new__(BeliefPropagationCache)
        c.qf = QuadraticFormNetwork(self.tensornetwork())
        c.ctx, c.lib, c.h = self.ctx, self.lib, h
        self.ctx._nets.add(c)
        c.storage, c.storage_name = self.storage, self.storage_name
        c._host_psi = c.qf.ket
        c._psi_dirty = False
        c.last_update = self.last_update
        c.partitioning = self.partitioning
        return c

    # -- structure -------------------------------------------------------------
    @property
    def graph(self):
        return self.qf.ket.graph

    def tensornetwork(self) -> ITensorNetwork:
        """The ket network (site tensors as held on the device)."""
        if self._psi_dirty:
            n = int(self.lib.itnbp_net_psi_numel(self.h))
            dt = {L.F64: np.float64, L.F32: np.float32, L.BF16: np.uint16}[self.storage]
            flat = np.empty(n, dtype=dt)
            self.ctx.check(self.lib.itnbp_get_sites(self.h, L.vptr(flat)))
            # bond dimensions as on the device (apply changes them)
            g = self.graph
            dims = np.empty(max(g.ne, 1), dtype=np.int32)
            self.ctx.check(self.lib.itnbp_get_bond_dims(self.h, L.iptr(dims)))
            shapes = {v: (self._host_psi.phys_dim(v),) + tuple(int(dims[g.edge_id(v, w)]) for w in g.incident(v))
                      for v in g.vertices}
            self._host_psi = _unflat_sites(self._host_psi, flat, self.storage, shapes)
            self.qf = QuadraticFormNetwork(self._host_psi)
            self._psi_dirty = False
        return self._host_psi

    def __getitem__(self, v):
        return self.tensornetwork()[v]

    def __setitem__(self, v, t):
        """bpc[v] = t: replaces the ket tensor of v (the bra follows as its conjugate)."""
        psi = self.tensornetwork()
        psi[v] = t
        i = self.graph.index[v]
        arr = np.asarray(t).reshape(-1, order="F")
        if self.storage == L.F64:
            buf = np.ascontiguousarray(arr, dtype=np.float64)
        elif self.storage == L.F32:
            buf = np.ascontiguousarray(arr, dtype=np.float32)
        else:
            back, buf = bf16_round(arr)
            psi[v] = back.reshape(psi[v].shape, order="F")
        self.ctx.check(self.lib.itnbp_set_site(self.h, i, L.vptr(buf)))

    def scalartype(self):
        return {L.F64: np.float64, L.F32: np.float32, L.BF16: np.float64}[self.storage]

    # -- messages --------------------------------------------------------------
    def _mid(self, e):
        a, b = e
        return self.graph.message_id(a, b)

    def has_message(self, e):
        return bool(self.lib.itnbp_message_is_set(self.h, self._mid(e)))

    def message(self, e):
        """message(bpc, e) (abstractbeliefpropagationcache.jl:173-175): M[b_ket, b_bra]."""
        a, b = e
        chi = self.tensornetwork().bond_dim(a, b)
        out = np.empty(chi * chi, dtype=np.float64)
        self.ctx.check(self.lib.itnbp_get_message(self.h, self._mid(e), L.dptr(out)))
        return out.reshape(chi, chi, order="F").astype(self.scalartype())

    def messages(self):
        """Dictionary of every set message."""
        n = int(self.lib.itnbp_net_msg_numel(self.h))
        flat = np.empty(max(n, 1), dtype=np.float64)
        self.ctx.check(self.lib.itnbp_get_messages(self.h, L.dptr(flat)))
        out = {}
        off = 0
        g = self.graph
        ket = self.tensornetwork()
        for m in range(2 * g.ne):
            a, b = g.message_edge(m)
            chi = ket.bond_dim(a, b)
            if self.lib.itnbp_message_is_set(self.h, m):
                out[(a, b)] = flat[off:off + chi * chi].reshape(chi, chi, order="F").astype(self.scalartype())
            off += chi * chi
        return out

    def set_message_(self, e, m):
        a, b = e
        chi = self.tensornetwork().bond_dim(a, b)
        m = np.asarray(m, dtype=np.float64)
        if m.shape != (chi, chi):
            raise ValueError(f"message shape {m.shape} != {(chi, chi)}")
        buf = np.ascontiguousarray(m.reshape(-1, order="F"))
        self.ctx.check(self.lib.itnbp_set_message(self.h, self._mid(e), L.dptr(buf)))
        return self

    def set_messages_(self, msgs):
        for e, m in msgs.items():
            self.set_message_(e, m)
        return self

    def set_message(self, e, m):
        return self.copy().set_message_(e, m)

    def set_messages(self, msgs):
        return self.copy().set_messages_(msgs)

    def update_factor(self, v, t):
        c = self.copy()
        c[v] = t
        return c

    def default_bp_maxiter(self):
        """1 on trees, nothing otherwise (abstractbeliefpropagationcache.jl:38-41)."""
        return 1 if self.graph.is_tree() else None


# ----------------------------------------------------------------------------
def identity_messages(qf):
    """delta(b_bra, b_ket) on every directed edge (quadraticformnetwork.jl:96-124)."""
    psi = qf.ket if isinstance(qf, QuadraticFormNetwork) else qf
    g = psi.graph
    out = {}
    for (a, b) in g.edges:
        chi = psi.bond_dim(a, b)
        out[(a, b)] = np.eye(chi)
        out[(b, a)] = np.eye(chi)
    return out


def initialize_cache(tn, alg="bp", *, messages=None, storage=None, ctx=None, partitioned_vertices=None):
    """initialize_cache(Algorithm"bp", ::QuadraticFormNetwork) (initialize_cache.jl:14-29):
    empty messages on trees, identity messages on loopy graphs.
    ``partitioned_vertices`` (BeliefPropagationCache(tn; partitioned_vertices),
    beliefpropagationcache.jl:20-35) groups vertices into partitions: the cache
    is then built over the partition network (see itnbp.partition)."""
    if alg != "bp":
        raise ValueError(f"unknown cache algorithm {alg!r}")
    if partitioned_vertices is not None:
        from .partition import PartitionedNetwork
        psi = tn.ket if isinstance(tn, QuadraticFormNetwork) else tn
        pn = PartitionedNetwork(psi, partitioned_vertices, ctx=ctx)
        bpc = initialize_cache(pn.network, messages=messages, storage=storage, ctx=ctx)
        bpc.partitioning = pn
        return bpc
    qf = tn if isinstance(tn, QuadraticFormNetwork) else QuadraticFormNetwork(tn)
    if np.iscomplexobj(np.zeros(0, dtype=qf.ket.dtype)):
        # complex states: BP on the norm network by the two-layer (bra, ket) kernels
        from .layer import norm_cache
        if storage not in (None, "f64"):
            raise ValueError("complex site tensors are held in complex FP64")
        return norm_cache(qf.ket, messages=messages, ctx=ctx)
    bpc = BeliefPropagationCache(qf, storage=storage, ctx=ctx)
    if messages is None:
        if not qf.is_tree():
            bpc.ctx.check(bpc.lib.itnbp_init_identity_messages(bpc.h))
    else:
        bpc.set_messages_(messages)
    return bpc


def default_edge_sequence(bpc):
    """Forest-cover sequential schedule (edge_sequences.jl:9-47) as directed edges."""
    g = bpc.graph
    out = np.empty(2 * g.ne, dtype=np.int32)
    bpc.ctx.check(bpc.lib.itnbp_forest_cover_sequence(bpc.h, L.iptr(out)))
    return [g.message_edge(int(m)) for m in out]


def _directed_edge(g, x):
    """``x`` as a directed edge (a, b) of ``g``, or None when it is not one (then
    it is a group of edges)."""
    try:
        if len(x) == 2 and x[0] in g.index and x[1] in g.index and g.has_edge(x[0], x[1]):
            return (x[0], x[1])
    except TypeError:
        pass
    return None


_GROUPED = "grouped"


def _schedule(bpc, edge_sequence):
    """(schedule, message ids) for the C ABI; (_GROUPED, groups) for edge groups
    with more than one edge, which the host runs group by group."""
    g = bpc.graph
    if edge_sequence is None or edge_sequence == "forest_cover":
        return L.SCHED_SEQUENTIAL, None
    if edge_sequence == "parallel":
        return L.SCHED_PARALLEL, None
    seq = list(edge_sequence)
    if seq and _directed_edge(g, seq[0]) is None:
        groups = []
        for grp in seq:
            edges = [_directed_edge(g, e) for e in grp]
            if not edges or any(e is None for e in edges):
                raise ValueError(f"edge group {grp!r} is not a non-empty list of directed edges of the graph")
            groups.append(edges)
        if any(len(grp) > 1 for grp in groups):
            return _GROUPED, groups
        ids = [g.message_id(*grp[0]) for grp in groups]
        return L.SCHED_PARALLEL, np.array(ids, dtype=np.int32)
    edges = [_directed_edge(g, e) for e in seq]
    if any(e is None for e in edges):
        raise ValueError("edge_sequence mixes edges and edge groups, or names an edge the graph does not have")
    ids = [g.message_id(*e) for e in edges]
    return L.SCHED_SEQUENTIAL, np.array(ids, dtype=np.int32)


def _update_grouped_(bpc, groups, maxiter, tol, normalize, precision):
    """update_iteration over edge groups (abstractbeliefpropagationcache.jl:294-308):
    every group is updated sequentially, starting from the messages the sweep
    started from, and the groups' results are merged; the convergence measure is
    the summed message_diff over the number of groups (:313-329)."""
    g = bpc.graph
    it = 0
    mean = 0.0
    for it in range(1, int(maxiter) + 1):
        new, diff = {}, 0.0
        for grp in groups:
            work = bpc.copy()
            o = L.UpdateOpts()
            o.schedule = L.SCHED_SEQUENTIAL
            o.maxiter = 1
            o.tol = 0.0
            o.res_tol = -1.0
            o.precision = PRECISIONS[precision]
            o.switch_tol = 0.0
            o.normalize = int(normalize)
            o.verbose = 0
            ids = np.array([g.message_id(*e) for e in grp], dtype=np.int32)
            o.nseq = len(ids)
            o.seq = L.iptr(ids)
            r = L.UpdateResult()
            bpc.ctx.check(bpc.lib.itnbp_update(work.h, C.byref(o), C.byref(r)))
            diff += r.mean_diff * len(grp)
            for e in grp:
                new[e] = work.message(e)
            del work
        bpc.set_messages_(new)
        mean = diff / len(groups)
        if tol is not None and mean <= tol:
            break
    bpc.last_update = dict(iterations=it, iterations_low=0, mean_diff=mean, max_residual=float("nan"),
                           converged=tol is not None and mean <= tol, device_ms=0.0, path="grouped-host")
    return bpc


def update_(bpc, alg="bp", *, maxiter="default", tol=None, edge_sequence=None, verbose=False,
            message_update_alg=None, precision="f64", res_tol=None, switch_tol=1e-4):
    """In-place update(bpc; alg="bp", maxiter, tol, edge_sequence, verbose,
    message_update_alg) (abstractbeliefpropagationcache.jl:313-337).

    ``edge_sequence``: None / "forest_cover" (sequential, the default), "parallel",
    a list of directed edges (sequential) or a list of edge groups (parallel).
    Extensions: ``res_tol`` stops on max ||new-old||/||new||; ``precision`` is
    "f64", "f32" (16-bit MFMA sweeps where the chi = 32 pipelines apply, FP32
    elsewhere), "bf16x3" (compensated 16-bit MFMA: operands split into three
    BF16 terms, FP32 intermediates), "bf16_f64" / "bf16x3_f64" (those sweeps
    until the residual drops below ``switch_tol`` or stagnates, then FP64), or
    "bf16_bf16x3" (the "f32" sweeps until then, followed by "bf16x3"), or
    "bf16_bf16x3_f64" ("f32" sweeps to residual 1e-2, "bf16x3" to ``switch_tol``,
    then FP64); "f16x2" is the compensated arithmetic with two FP16 terms per
    operand (about 22 bits) on the chi = 32 eigenbasis pipeline, and
    "bf16_f16x2" the "f32" sweeps until ``switch_tol`` followed by "f16x2"."""
    if alg != "bp":
        raise ValueError(f"unknown update algorithm {alg!r}")
    if getattr(bpc, "layered", False):
        if precision != "f64" or message_update_alg is not None:
            raise ValueError("the two-layer engine runs the FP64 'contract' update only")
        if edge_sequence is not None and edge_sequence not in ("forest_cover", "parallel") and \
                isinstance(edge_sequence[0], list):
            raise ValueError("edge groups are not available on the two-layer engine")
        return bpc.update_(maxiter=maxiter, tol=tol, res_tol=res_tol, edge_sequence=edge_sequence,
                           verbose=verbose)
    if isinstance(maxiter, str):
        maxiter = bpc.default_bp_maxiter()
    if maxiter is None:
        raise ValueError("You need to specify a number of iterations for BP!")
    normalize = True
    if message_update_alg is not None:
        mua = dict(message_update_alg)
        if mua.get("alg", "contract") != "contract":
            raise ValueError("only the 'contract' message update is available")
        normalize = bool(mua.get("normalize", True))
    sched, seq = _schedule(bpc, edge_sequence)
    if sched == _GROUPED:
        return _update_grouped_(bpc, seq, maxiter, tol, normalize, precision)
    o = L.UpdateOpts()
    o.schedule = sched
    o.maxiter = int(maxiter)
    o.tol = -1.0 if tol is None else float(tol)
    o.res_tol = -1.0 if res_tol is None else float(res_tol)
    o.precision = PRECISIONS[precision]
    o.switch_tol = float(switch_tol)
    o.normalize = int(normalize)
    o.verbose = int(bool(verbose))
    if seq is not None:
        o.nseq = len(seq)
        o.seq = L.iptr(seq)
    else:
        o.nseq = 0
        o.seq = None
    r = L.UpdateResult()
    bpc.ctx.check(bpc.lib.itnbp_update(bpc.h, C.byref(o), C.byref(r)))
    bpc.last_update = dict(iterations=r.iterations, iterations_low=r.iterations_low,
                           mean_diff=r.mean_diff, max_residual=r.max_residual,
                           converged=bool(r.converged), device_ms=r.device_ms,
                           path=bpc.ctx.last_path())
    return bpc


def update(bpc, alg="bp", **kwargs):
    """Functional update: returns a new cache (abstractbeliefpropagationcache.jl:331-337)."""
    return update_(bpc.copy(), alg, **kwargs)


def updated_message(bpc, e, *, normalize=True):
    """updated_message(bpc, e) (abstractbeliefpropagationcache.jl:225-239, 254-261)."""
    a, b = e
    chi = bpc.tensornetwork().bond_dim(a, b)
    ids = np.array([bpc.graph.message_id(a, b)], dtype=np.int32)
    out = np.empty(chi * chi, dtype=np.float64)
    bpc.ctx.check(bpc.lib.itnbp_updated_messages(bpc.h, 1, L.iptr(ids), int(normalize), L.dptr(out)))
    return out.reshape(chi, chi, order="F").astype(bpc.scalartype())


def message_diff(a, b):
    """1 - |<a/|a|, b/|b|>|^2 (abstractbeliefpropagationcache.jl:32-36)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    f = abs(np.vdot(a / np.linalg.norm(a), b / np.linalg.norm(b))) ** 2
    return 1.0 - f


# ----------------------------------------------------------------------------
def environment(bpc, verts):
    """environment(bpc, verts) for whole partitions (beliefpropagationcache.jl:100-105):
    the incoming messages into the set ``verts`` from outside it, as a dict
    {(w, v): M_{w->v}}.  For one or two neighbouring vertices this is the
    product environment ``apply`` needs."""
    g = bpc.graph
    vs = set(verts)
    out = {}
    for v in verts:
        for w in g.incident(v):
            if w not in vs:
                out[(w, v)] = bpc.message((w, v))
    return out


def rdm(bpc, verts):
    """BP reduced density matrix of one vertex (d x d, rho[s, s']) or of two
    neighbouring vertices ((d_u d_v)^2, first site fastest), unnormalised."""
    g = bpc.graph
    if len(verts) == 1:
        v = verts[0]
        d = bpc.qf.ket.phys_dim(v)
        out = np.empty(d * d, dtype=np.float64)
        ids = np.array([g.index[v]], dtype=np.int32)
        bpc.ctx.check(bpc.lib.itnbp_rdm1(bpc.h, 1, L.iptr(ids), L.dptr(out)))
        return out.reshape(d, d, order="F")
    if len(verts) == 2:
        u, v = verts
        e = g.edge_id(u, v)
        du, dv = bpc.qf.ket.phys_dim(g.edges[e][0]), bpc.qf.ket.phys_dim(g.edges[e][1])
        out = np.empty((du * dv) ** 2, dtype=np.float64)
        ids = np.array([e], dtype=np.int32)
        bpc.ctx.check(bpc.lib.itnbp_rdm2(bpc.h, 1, L.iptr(ids), L.dptr(out)))
        r = out.reshape(du * dv, du * dv, order="F")
        if g.edges[e] != (u, v):  # reorder to (u, v) with u fastest
            r = r.reshape(du, dv, du, dv, order="F").transpose(1, 0, 3, 2).reshape(du * dv, du * dv, order="F")
        return r
    raise ValueError("rdm supports one vertex or one neighbouring pair")


def expect_on_cache(bpc, op, vertices=None):
    """<O_v> = tr(O rho_v)/tr(rho_v) for every vertex (expect.jl:5-19).  On a
    cache over generalised partitions, rho_v is the partial trace of v's
    partition RDM and ``vertices`` are vertices of the original network."""
    if getattr(bpc, "layered", False):
        return bpc.expect(op, vertices)
    if bpc.partitioning is not None:
        pn = bpc.partitioning
        vertices = list(pn.psi.graph.vertices) if vertices is None else list(vertices)
        rho_p, out = {}, {}
        for v in vertices:
            p = pn.owner[v]
            if p not in rho_p:
                rho_p[p] = rdm(bpc, [p])
            r = pn.member_rdm(rho_p[p], v)
            o = op_matrix(op, pn.psi.phys_dim(v))
            out[v] = float(np.trace(o @ r) / np.trace(r))
        return out
    g = bpc.graph
    vertices = list(g.vertices) if vertices is None else list(vertices)
    ids = np.array([g.index[v] for v in vertices], dtype=np.int32)
    mats = np.concatenate([op_matrix(op, bpc.qf.ket.phys_dim(v)).reshape(-1, order="F") for v in vertices])
    out = np.empty(len(vertices), dtype=np.float64)
    bpc.ctx.check(bpc.lib.itnbp_expect1(bpc.h, len(vertices), L.iptr(ids), L.dptr(mats), L.dptr(out)))
    return dict(zip(vertices, out))


def expect2_on_cache(bpc, op2, edges=None):
    """<O_uv> for neighbouring pairs from the BP pair environment (the two-site RDM
    pattern of test/test_belief_propagation.jl:64-91); ``op2`` is a (d_u d_v)^2
    matrix O[(s_u', s_v'), (s_u, s_v)] with s_u fastest, or a callable (u, v) -> matrix."""
    g = bpc.graph
    edges = list(g.edges) if edges is None else list(edges)
    ids, mats = [], []
    for (u, v) in edges:
        e = g.edge_id(u, v)
        m = op2(u, v) if callable(op2) else np.asarray(op2, dtype=np.float64)
        if g.edges[e] != (u, v):
            du, dv = bpc.qf.ket.phys_dim(u), bpc.qf.ket.phys_dim(v)
            m = m.reshape(du, dv, du, dv, order="F").transpose(1, 0, 3, 2).reshape(du * dv, du * dv, order="F")
        ids.append(e)
        mats.append(np.asarray(m, dtype=np.float64).reshape(-1, order="F"))
    ids = np.array(ids, dtype=np.int32)
    mats = np.concatenate(mats)
    out = np.empty(len(edges), dtype=np.float64)
    bpc.ctx.check(bpc.lib.itnbp_expect2(bpc.h, len(edges), L.iptr(ids), L.dptr(mats), L.dptr(out)))
    return dict(zip(edges, out))


def expect(psi, op, vertices=None, *, alg="bp", cache=None, update_cache=None,
           cache_update_kwargs=None, cache_construction_kwargs=None):
    """expect(psi, op; alg="bp", cache!, update_cache, cache_update_kwargs,
    cache_construction_kwargs) (expect.jl:21-41, 102-109).  ``cache`` is a
    one-element list used like the reference's ``cache!::Ref``."""
    if alg != "bp":
        raise ValueError("only alg='bp' runs on the device; exact contraction is not part of this library")
    if update_cache is None:
        update_cache = cache is None
    if cache is None:
        cache = [initialize_cache(psi, **(cache_construction_kwargs or {}))]
    if update_cache:
        cache[0] = update_(cache[0], **(cache_update_kwargs or {}))
    return expect_on_cache(cache[0], op, vertices)


def region_scalars(bpc):
    g = bpc.graph
    vs = np.empty(g.nv, dtype=np.float64)
    es = np.empty(max(g.ne, 1), dtype=np.float64)
    bpc.ctx.check(bpc.lib.itnbp_region_scalars(bpc.h, L.dptr(vs), L.dptr(es)))
    return vs, es[:g.ne]


def vertex_scalars(bpc):
    return region_scalars(bpc)[0]


def edge_scalars(bpc):
    return region_scalars(bpc)[1]


def logscalar(bpc):
    """logscalar(bpc) (abstractbeliefpropagationcache.jl:397-408)."""
    if getattr(bpc, "layered", False):
        return bpc.logscalar()
    re, im, ninf = C.c_double(), C.c_double(), C.c_int()
    bpc.ctx.check(bpc.lib.itnbp_logscalar(bpc.h, C.byref(re), C.byref(im), C.byref(ninf)))
    if ninf.value:
        return -np.inf
    return complex(re.value, im.value) if im.value != 0 else re.value


def scalar(bpc):
    return np.exp(logscalar(bpc))


def rescale_(bpc):
    """rescale(bpc) = rescale_messages then rescale_partitions of the ket and bra
    vertices (abstractbeliefpropagationcache.jl:391-395)."""
    bpc.ctx.check(bpc.lib.itnbp_normalize(bpc.h))
    bpc._psi_dirty = True
    return bpc


def rescale(bpc):
    return rescale_(bpc.copy())


def normalize(psi, alg="bp", *, cache=None, update_cache=None, cache_update_kwargs=None,
              cache_construction_kwargs=None):
    """normalize(psi; alg="bp") (normalize.jl:125-142): returns the rescaled ket network."""
    if alg != "bp":
        raise ValueError("only alg='bp' runs on the device")
    if update_cache is None:
        update_cache = cache is None
    if cache is None:
        cache = [initialize_cache(psi, **(cache_construction_kwargs or {}))]
    if update_cache:
        cache[0] = update_(cache[0], **(cache_update_kwargs or {}))
    cache[0] = rescale_(cache[0])
    return cache[0].tensornetwork()


def energy_and_gradient(psi, nsweeps, ops1=None, ops2=None, *, checkpoint=1, ctx=None):
    """BP energy E = sum_v <ops1[v]> + sum_e <ops2[e]> after ``nsweeps`` Jacobi sweeps
    from identity messages, and dE/dpsi_v for every site tensor, by the library's
    hand-written adjoint kernels (sweeps checkpointed every ``checkpoint``).  This is
    the pullback the Julia shim wraps as a ChainRulesCore rrule; the reference has
    no autodiff.  ``ops1``: {v: d x d} or callable; ``ops2``: {(u, v): (d_u d_v)^2} or
    callable (u, v) -> matrix (first site fastest).  Returns (E, {v: grad})."""
    bpc = BeliefPropagationCache(QuadraticFormNetwork(psi), storage="f64", ctx=ctx)
    g = psi.graph
    o1 = o2 = None
    if ops1 is not None:
        o1 = np.concatenate([np.asarray(ops1(v) if callable(ops1) else ops1[v], dtype=np.float64).reshape(-1, order="F")
                             for v in g.vertices])
    if ops2 is not None:
        mats = []
        for (u, v) in g.edges:
            m = np.asarray(ops2(u, v) if callable(ops2) else ops2[(u, v)], dtype=np.float64)
            mats.append(m.reshape(-1, order="F"))
        o2 = np.concatenate(mats) if mats else None
    E = C.c_double()
    grad = np.empty(int(bpc.lib.itnbp_net_psi_numel(bpc.h)), dtype=np.float64)
    bpc.ctx.check(bpc.lib.itnbp_energy_grad(bpc.h, int(nsweeps), int(checkpoint),
                                            L.dptr(o1) if o1 is not None else None,
                                            L.dptr(o2) if o2 is not None else None,
                                            C.byref(E), L.dptr(grad)))
    out, off = {}, 0
    for v in g.vertices:
        n = psi[v].size
        out[v] = grad[off:off + n].reshape(psi[v].shape, order="F")
        off += n
    return E.value, out


def update_batched_(caches, *, maxiter, tol=None, res_tol=None, precision="f64", normalize=True):
    """Update many independent caches (any lattices and bond dimensions, one storage
    dtype) together: every message of every instance in the same launches of a
    Jacobi sweep; converged instances drop out.  In place; returns the per-cache
    results (each also in cache.last_update)."""
    if not caches:
        return []
    if maxiter is None:
        raise ValueError("You need to specify a number of iterations for BP!")
    n = len(caches)
    hs = (C.c_void_p * n)(*[c.h.value for c in caches])
    o = L.UpdateOpts()
    o.schedule = L.SCHED_PARALLEL
    o.maxiter = int(maxiter)
    o.tol = -1.0 if tol is None else float(tol)
    o.res_tol = -1.0 if res_tol is None else float(res_tol)
    o.precision = PRECISIONS[precision]
    o.switch_tol = 0.0
    o.normalize = int(normalize)
    o.nseq = 0
    o.seq = None
    o.verbose = 0
    res = (L.UpdateResult * n)()
    ctx = caches[0].ctx
    ctx.check(ctx.lib.itnbp_update_batched(n, hs, C.byref(o), res))
    out = []
    for c, r in zip(caches, res):
        c.last_update = dict(iterations=r.iterations, iterations_low=r.iterations_low, mean_diff=r.mean_diff,
                             max_residual=r.max_residual, converged=bool(r.converged), device_ms=r.device_ms,
                             path=ctx.last_path())
        out.append(c.last_update)
    return out


def sweeps_vjp(psi, nsweeps, mbar, *, checkpoint=1, ctx=None):
    """Pullback through ``nsweeps`` Jacobi sweeps from identity messages: given the
    cotangents ``mbar`` ({directed edge: matrix}; absent messages count as zero) of
    the final messages, the gradient with respect to every site tensor
    ({v: array}).  With ``apply_gates_vjp`` it gives the gradient of a TEBD step.
    The reference has no autodiff."""
    bpc = BeliefPropagationCache(QuadraticFormNetwork(psi), storage="f64", ctx=ctx)
    g = psi.graph
    mb = np.zeros(int(bpc.lib.itnbp_net_msg_numel(bpc.h)))
    off = 0
    for mid in range(2 * g.ne):
        a, b = g.message_edge(mid)
        chi = psi.bond_dim(a, b)
        if (a, b) in mbar:
            mb[off:off + chi * chi] = np.asarray(mbar[(a, b)], dtype=np.float64).reshape(-1, order="F")
        off += chi * chi
    grad = np.empty(int(bpc.lib.itnbp_net_psi_numel(bpc.h)), dtype=np.float64)
    bpc.ctx.check(bpc.lib.itnbp_sweeps_vjp(bpc.h, int(nsweeps), int(checkpoint), L.dptr(mb), L.dptr(grad)))
    out, off = {}, 0
    for v in g.vertices:
        n = psi[v].size
        out[v] = grad[off:off + n].reshape(psi[v].shape, order="F")
        off += n
    return out
