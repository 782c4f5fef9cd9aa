"""Two-layer BP on the device: inner products with alg="bp" and BP on the norm
network of a complex state.

Mirrors (paths relative to the ITensorNetworks.jl repository):
  inner / loginner (alg="bp")          src/inner.jl:100-160
  inner_network                        src/abstractitensornetwork.jl:509-516
  LinearFormNetwork                    src/formnetworks/linearformnetwork.jl:15-25
  BilinearFormNetwork                  src/formnetworks/bilinearformnetwork.jl:23-42
  initialize_cache (no message default) src/initialize_cache.jl:10-12
  update / logscalar / region_scalar   src/caches/abstractbeliefpropagationcache.jl:294-337, 397-412,
                                       src/caches/beliefpropagationcache.jl:107-119

Every contraction runs in libitnbp.so (k_layer.hip); the host only marshals.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .network import ITensorNetwork


def _is_complex(*arrays):
    return any(np.iscomplexobj(a) for a in arrays)


def _flat(tensors, vertices, cplx):
    parts = [np.asarray(tensors[v]).reshape(-1, order="F") for v in vertices]
    flat = np.concatenate(parts) if parts else np.zeros(0)
    if cplx:
        return np.ascontiguousarray(flat, dtype=np.complex128).view(np.float64)
    return np.ascontiguousarray(flat, dtype=np.float64)


def _bond_dims(graph, tensors, first_bond_axis=1):
    return np.array([np.asarray(tensors[a]).shape[first_bond_axis + graph.axis(a, b) - 1] for (a, b) in graph.edges],
                    dtype=np.int32)


class LayeredBPCache:
    """BP cache of a two-layer network <bra|ket> held on one GPU: the
    LinearFormNetwork of inner(bra, ket), of inner(bra, A, ket) with the
    operator folded into the ket (``operator``), or with bra = ket the norm
    network of a (complex) state.  Site tensors ``[s, bonds in incident
    order]``; an operator network's tensors are ``A_v[s', s, bonds]``.
    Messages are ``M[b_ket, b_bra]`` and have no default (initialize_cache.jl:10-12);
    ``messages="identity"`` sets delta(b_ket, b_bra) where the bond dimensions agree."""

    layered = True

    def __init__(self, bra: ITensorNetwork, ket: ITensorNetwork, *, operator=None, messages=None, ctx=None):
        g = ket.graph
        if bra.graph.vertices != g.vertices or bra.graph.edges != g.edges:
            raise ValueError("bra and ket must be networks on the same graph")
        for v in g.vertices:
            if bra.graph.incident(v) != g.incident(v):
                raise ValueError("bra and ket must order the bonds of every site alike")
        self.graph = g
        self.bra, self.ket, self.operator = bra, ket, operator
        self.ctx = ctx or L.default_context()
        self.lib = self.ctx.lib
        vs = g.vertices
        tens = [ket[v] for v in vs] + [bra[v] for v in vs] + ([operator[v] for v in vs] if operator else [])
        self.cplx = _is_complex(*tens)
        self.dtype = np.complex128 if self.cplx else np.float64
        edges, inc_ptr, inc_edge = g.csr()
        kchi = _bond_dims(g, {v: ket[v] for v in vs})
        if operator is not None:
            achi = _bond_dims(g, {v: operator[v] for v in vs}, first_bond_axis=2)
            kchi_net = kchi * achi
        else:
            kchi_net = kchi
        bchi = _bond_dims(g, {v: bra[v] for v in vs})
        phys = np.array([np.asarray(ket[v]).shape[0] for v in vs], dtype=np.int32)
        for v in vs:
            if np.asarray(bra[v]).shape[0] != phys[g.index[v]]:
                raise ValueError(f"physical dimensions of bra and ket differ at {v}")
        self.kchi, self.bchi, self.phys = kchi_net, bchi, phys
        h = C.c_void_p()
        ne = g.ne
        self.ctx.check(self.lib.itnbp_lnet_create(
            self.ctx.h, g.nv, ne, L.iptr(edges) if ne else None, L.iptr(inc_ptr), L.iptr(inc_edge) if ne else None,
            L.iptr(kchi_net) if ne else None, L.iptr(bchi) if ne else None, L.iptr(phys), int(self.cplx),
            C.byref(h)))
        self.h = h
        self.ctx._nets.add(self)
        bra_flat = _flat(bra, vs, self.cplx)
        if operator is None:
            self.ctx.check(self.lib.itnbp_lnet_set_tensors(self.h, L.dptr(_flat(ket, vs, self.cplx)), L.dptr(bra_flat)))
        else:
            nk = int(self.lib.itnbp_lnet_ket_numel(self.h))
            zeros = np.zeros(nk * (2 if self.cplx else 1))
            self.ctx.check(self.lib.itnbp_lnet_set_tensors(self.h, L.dptr(zeros), L.dptr(bra_flat)))
            self.ctx.check(self.lib.itnbp_lnet_fold_operator(
                self.h, L.dptr(_flat(ket, vs, self.cplx)), L.iptr(kchi) if ne else None,
                L.dptr(_flat(operator, vs, self.cplx)), L.iptr(achi) if ne else None))
        self.last_update = None
        if messages == "identity":
            self.ctx.check(self.lib.itnbp_lnet_init_identity_messages(self.h))
        elif messages:
            for e, m in messages.items():
                self.set_message_(e, m)

    def _release(self):
        if getattr(self, "h", None):
            self.lib.itnbp_lnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def scalartype(self):
        return self.dtype

    def copy(self):
        """A new device network with the same tensors and messages
        (beliefpropagationcache.jl:43-47)."""
        c = LayeredBPCache(self.bra, self.ket, operator=self.operator, ctx=self.ctx)
        for e, m in self.messages().items():
            c.set_message_(e, m)
        c.last_update = self.last_update
        return c

    # -- messages ------------------------------------------------------------
    def _msg_shape(self, a, b):
        e = self.graph.edge_id(a, b)
        return int(self.kchi[e]), int(self.bchi[e])

    def message(self, e):
        a, b = e
        m = self.graph.message_id(a, b)
        shape = self._msg_shape(a, b)
        buf = np.empty(int(np.prod(shape)), dtype=self.dtype)
        self.ctx.check(self.lib.itnbp_lnet_get_message(self.h, m, L.dptr(buf.view(np.float64))))
        return buf.reshape(shape, order="F")

    def set_message_(self, e, M):
        a, b = e
        m = self.graph.message_id(a, b)
        shape = self._msg_shape(a, b)
        M = np.asarray(M)
        if M.shape != shape:
            raise ValueError(f"message {e} must have shape {shape}, got {M.shape}")
        if np.iscomplexobj(M) and not self.cplx:
            raise ValueError("complex message on a real network")
        flat = np.ascontiguousarray(M.reshape(-1, order="F"), dtype=self.dtype)
        self.ctx.check(self.lib.itnbp_lnet_set_message(self.h, m, L.dptr(flat.view(np.float64))))
        return self

    def messages(self):
        g = self.graph
        out = {}
        for m in range(2 * g.ne):
            if self.lib.itnbp_lnet_message_is_set(self.h, m) == 1:
                e = g.message_edge(m)
                out[e] = self.message(e)
        return out

    def folded_ket(self):
        """The ket site tensors as held on the device (A psi with merged bonds
        when an operator was folded in)."""
        g = self.graph
        n = int(self.lib.itnbp_lnet_ket_numel(self.h))
        flat = np.empty(n, dtype=self.dtype)
        self.ctx.check(self.lib.itnbp_lnet_get_ket(self.h, L.dptr(flat.view(np.float64))))
        out, off = {}, 0
        for v in g.vertices:
            shape = (int(self.phys[g.index[v]]),) + tuple(int(self.kchi[g.edge_id(v, w)]) for w in g.incident(v))
            k = int(np.prod(shape))
            out[v] = flat[off:off + k].reshape(shape, order="F")
            off += k
        return ITensorNetwork(g, out)

    # -- BP ------------------------------------------------------------------
    def update_(self, *, maxiter="default", tol=None, res_tol=None, edge_sequence=None, normalize=True,
                verbose=False):
        """update(bpc; alg="bp", maxiter, tol, edge_sequence)
        (abstractbeliefpropagationcache.jl:294-337): maxiter defaults to 1 on a
        tree and is required otherwise (:38, :315-317); the default sequence is
        the forest cover (sequential)."""
        g = self.graph
        if maxiter == "default":
            maxiter = 1 if g.is_tree() else None
        if maxiter is None:
            raise ValueError("You need to specify a number of iterations for BP!")
        o = L.UpdateOpts()
        seq = None
        if edge_sequence is None or edge_sequence == "forest_cover":
            o.schedule = L.SCHED_SEQUENTIAL
        elif edge_sequence == "parallel":
            o.schedule = L.SCHED_PARALLEL
        else:
            o.schedule = L.SCHED_SEQUENTIAL
            seq = np.array([g.message_id(a, b) for (a, b) in edge_sequence], dtype=np.int32)
        o.maxiter = int(maxiter)
        o.tol = -1.0 if tol is None else float(tol)
        o.res_tol = -1.0 if res_tol is None else float(res_tol)
        o.precision = L.PREC_F64
        o.switch_tol = 0.0
        o.normalize = int(normalize)
        o.nseq = 0 if seq is None else len(seq)
        o.seq = None if seq is None else L.iptr(seq)
        o.verbose = int(verbose)
        r = L.UpdateResult()
        self.ctx.check(self.lib.itnbp_lnet_update(self.h, C.byref(o), C.byref(r)))
        self.last_update = dict(iterations=r.iterations, mean_diff=r.mean_diff, max_residual=r.max_residual,
                                converged=bool(r.converged), device_ms=r.device_ms, path=self.ctx.last_path())
        return self

    def region_scalars(self):
        g = self.graph
        vs = np.empty(2 * g.nv)
        es = np.empty(2 * max(g.ne, 1))
        self.ctx.check(self.lib.itnbp_lnet_region_scalars(self.h, L.dptr(vs), L.dptr(es)))
        vert = {v: complex(vs[2 * i], vs[2 * i + 1]) for i, v in enumerate(g.vertices)}
        edge = {e: complex(es[2 * i], es[2 * i + 1]) for i, e in enumerate(g.edges)}
        if not self.cplx:
            vert = {k: x.real for k, x in vert.items()}
            edge = {k: x.real for k, x in edge.items()}
        return vert, edge

    def logscalar(self):
        """sum(log.(vertex scalars)) - sum(log.(edge scalars)), complex logarithms
        (abstractbeliefpropagationcache.jl:397-412)."""
        re, im = C.c_double(), C.c_double()
        self.ctx.check(self.lib.itnbp_lnet_logscalar(self.h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value) if (self.cplx or im.value != 0) else re.value

    def scalar(self):
        return np.exp(self.logscalar())

    def rdm1(self, v):
        """R[s_ket, s_bra] of site v, unnormalised (environment contracted with
        the ket and the conjugated bra)."""
        d = int(self.phys[self.graph.index[v]])
        out = np.empty(d * d, dtype=self.dtype)
        verts = np.array([self.graph.index[v]], dtype=np.int32)
        self.ctx.check(self.lib.itnbp_lnet_rdm1(self.h, 1, L.iptr(verts), L.dptr(out.view(np.float64))))
        return out.reshape(d, d, order="F")

    def expect(self, op, vertices=None):
        """tr(O R_v) / tr(R_v) per vertex, O[s', s] (expect.jl:5-19)."""
        from .cache import op_matrix
        vertices = list(self.graph.vertices) if vertices is None else list(vertices)
        out = {}
        for v in vertices:
            R = self.rdm1(v)
            o = np.asarray(op) if not isinstance(op, str) else op_matrix(op, R.shape[0])
            val = np.trace(o @ R) / np.trace(R)
            out[v] = val if self.cplx else float(np.real(val))
        return out


def norm_cache(psi: ITensorNetwork, *, messages=None, ctx=None):
    """BP cache of <psi|psi> on the two-layer engine (real or complex psi):
    identity messages on loopy graphs, none on trees (initialize_cache.jl:14-29)."""
    if messages is None:
        messages = None if psi.graph.is_tree() else "identity"
    return LayeredBPCache(psi, psi, messages=messages, ctx=ctx)


def inner_network_cache(phi, *args, messages=None, ctx=None):
    """BP cache of inner_network(phi, psi) or inner_network(phi, A, psi)
    (abstractitensornetwork.jl:509-516)."""
    if len(args) == 1:
        return LayeredBPCache(phi, args[0], messages=messages, ctx=ctx)
    if len(args) == 2:
        A, psi = args
        return LayeredBPCache(phi, psi, operator=A, messages=messages, ctx=ctx)
    raise TypeError("inner(phi, psi) or inner(phi, A, psi)")


def loginner(phi, *args, alg="bp", cache_update_kwargs=None, messages=None, ctx=None):
    """loginner(phi, [A,] psi; alg="bp") (inner.jl:100-145): logscalar of the
    BP cache of the inner network."""
    if alg != "bp":
        raise ValueError("only alg='bp' runs on the device; exact contraction is not part of this library")
    c = inner_network_cache(phi, *args, messages=messages, ctx=ctx)
    c.update_(**(cache_update_kwargs or {}))
    return c.logscalar()


def inner(phi, *args, alg="bp", cache_update_kwargs=None, messages=None, ctx=None):
    """inner(phi, [A,] psi; alg="bp") (inner.jl:147-160): scalar of the BP cache
    of the inner network."""
    return np.exp(loginner(phi, *args, alg=alg, cache_update_kwargs=cache_update_kwargs, messages=messages, ctx=ctx))
