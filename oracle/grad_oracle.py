"""FP64 autograd oracle for the BP-energy gradient (test infrastructure only).

The reference has no autodiff (no ChainRulesCore / Zygote in its Project.toml),
so the gradient of a BP expectation value is pinned by (a) this PyTorch FP64
autograd restatement of the same unrolled computation and (b) central finite
differences of the forward energy.  The forward pass restates, in torch:

  * identity initial messages (src/formnetworks/quadraticformnetwork.jl:96-124);
  * ``nsweeps`` Jacobi sweeps of updated_message + normalisation
    (src/caches/abstractbeliefpropagationcache.jl:225-239, 294-308);
  * one-site <O_v> = tr(O rho_v)/tr(rho_v) (src/expect.jl:5-19) and two-site
    <O_uv> from the BP pair environment (test/test_belief_propagation.jl:64-91).

Conventions are those of ``bp_oracle``: psi[v] has shape (d, chi_e...) in the
graph's incident order; messages are M[b_ket, b_bra].
"""
from __future__ import annotations

import numpy as np
import torch


def _mode(X, M, ax):
    Y = torch.tensordot(X, M, dims=([ax], [0]))
    return torch.movedim(Y, -1, ax)


def _open(g, psi, msgs, u, open_axes):
    """Open contraction of vertex u (bp_oracle.open_contraction in torch): a
    (rows x rows) matrix whose row index is the column-major flattening of the
    open axes of the ket, the column index that of the bra."""
    t = psi[u]
    phi = t
    for a, (w, _) in enumerate(g.incident[u], start=1):
        if a in open_axes:
            continue
        phi = _mode(phi, msgs[(w, u)], a)
    rest = [a for a in range(t.dim()) if a not in open_axes]
    T = torch.tensordot(phi, t, dims=(rest, rest))
    if not open_axes:
        return T.reshape(1, 1)
    nd = len(open_axes)
    rows = int(np.prod([t.shape[a] for a in sorted(open_axes)]))
    perm = list(range(nd - 1, -1, -1)) + list(range(2 * nd - 1, nd - 1, -1))
    return T.permute(*perm).reshape(rows, rows)


def rdm2_torch(g, psi, msgs, u, v):
    """rho[(su, sv), (su', sv')] (su fastest) of the pair (u, v)."""
    du, dv = psi[u].shape[0], psi[v].shape[0]
    chi = psi[u].shape[g.axis(u, v)]
    Au = _open(g, psi, msgs, u, [0, g.axis(u, v)]).reshape(chi, du, chi, du).permute(1, 0, 3, 2)
    Av = _open(g, psi, msgs, v, [0, g.axis(v, u)]).reshape(chi, dv, chi, dv).permute(1, 0, 3, 2)
    rho = torch.einsum("ibjc,kblc->ikjl", Au, Av)  # [su, sv, su', sv']
    return rho.permute(1, 0, 3, 2).reshape(du * dv, du * dv)


def bp_energy(g, psi, nsweeps, ops1=None, ops2=None):
    """E = sum_v <ops1[v]> + sum_e <ops2[e]> after nsweeps Jacobi sweeps from
    identity messages; psi is a list of torch FP64 tensors."""
    msgs = {}
    for (u, v) in g.edges:
        chi = psi[u].shape[g.axis(u, v)]
        eye = torch.eye(chi, dtype=torch.float64)
        msgs[(u, v)] = eye
        msgs[(v, u)] = eye
    directed = g.directed_edges()
    for _ in range(nsweeps):
        new = {}
        for (u, v) in directed:
            T = _open(g, psi, msgs, u, [g.axis(u, v)])
            n = torch.linalg.norm(T)
            new[(u, v)] = T / n
        msgs = new
    E = torch.zeros((), dtype=torch.float64)
    if ops1 is not None:
        for u in range(g.nv):
            rho = _open(g, psi, msgs, u, [0])
            O = torch.as_tensor(ops1[u], dtype=torch.float64)
            E = E + torch.sum(O.T * rho) / torch.trace(rho)
    if ops2 is not None:
        for k, (u, v) in enumerate(g.edges):
            rho = rdm2_torch(g, psi, msgs, u, v)
            O = torch.as_tensor(ops2[k], dtype=torch.float64)
            E = E + torch.sum(O.T * rho) / torch.trace(rho)
    return E


def energy_and_grad(g, psi_np, nsweeps, ops1=None, ops2=None):
    psi = [torch.tensor(np.asarray(p, dtype=np.float64), requires_grad=True) for p in psi_np]
    E = bp_energy(g, psi, nsweeps, ops1, ops2)
    E.backward()
    return float(E.detach()), [p.grad.detach().numpy().copy() for p in psi]


# ----------------------------------------------------------------------------
# simple_update_bp in torch (the adjoint of the gate application)
# ----------------------------------------------------------------------------
def _map_eigvals_torch(f, A, cutoff=10 * 2.220446049250313e-16):
    """map_eigvals (apply.jl:9-25) in torch: eigenvalues dropped by the same rule
    as bp_oracle.map_eigvals (a constant selection), f on the kept ones."""
    from oracle import bp_oracle as O
    A = 0.5 * (A + A.T)
    w, U = torch.linalg.eigh(A)
    order = torch.argsort(-torch.abs(w).detach(), stable=True)
    w, U = w[order], U[:, order]
    keep, _ = O.truncation_keep(np.abs(w.detach().numpy()), cutoff)
    w, U = w[:keep], U[:, :keep]
    return (U * f(w)) @ U.T


def simple_update_bp_torch(G, psi_u, psi_v, ax_u, ax_v, envs_u, envs_v, maxdim=None, cutoff=None,
                           normalize=False):
    """bp_oracle.simple_update_bp (apply.jl:33-95, :134-137) restated in torch
    FP64 so that autograd gives its vector-Jacobian product: eigh-based sqrt and
    inverse sqrt of the environments, reduced QR of both sides, SVD of Theta
    truncated by the same (constant) rule, the same sign gauge of the new bond
    (constant signs) and the optional normalisation.  All arguments torch
    tensors except the axes; returns (psi_u', psi_v')."""
    from oracle import bp_oracle as O
    du, dv = psi_u.shape[0], psi_v.shape[0]

    def reduce_side(t, ax, envs):
        S = {a: _map_eigvals_torch(torch.sqrt, M) for a, M in envs.items()}
        Sinv = {a: _map_eigvals_torch(lambda x: 1.0 / torch.sqrt(x), M) for a, M in envs.items()}
        tt = t
        for a, s in S.items():
            tt = _mode(tt, s, a)
        ext = [a for a in range(1, t.dim()) if a != ax]
        tp = tt.permute(*(ext + [0, ax]))
        E = int(np.prod([t.shape[a] for a in ext])) if ext else 1
        # column-major reshape: reverse the axes, reshape row-major, transpose
        Xm = tp.permute(*reversed(range(tp.dim()))).reshape(t.shape[0] * t.shape[ax], E).T
        Q, R = torch.linalg.qr(Xm, mode="reduced")
        r = Q.shape[1]
        R3 = R.T.reshape(t.shape[ax], t.shape[0], r).permute(2, 1, 0)  # [r, s, b]
        return Q, R3, ext, Sinv

    Qu, Ru, ext_u, Sinv_u = reduce_side(psi_u, ax_u, envs_u)
    Qv, Rv, ext_v, Sinv_v = reduce_side(psi_v, ax_v, envs_v)
    ru, rv = Ru.shape[0], Rv.shape[0]
    G4 = G.T.reshape(dv, du, dv, du).permute(3, 2, 1, 0)  # column-major [su', sv', su, sv]
    RR = torch.einsum("aib,cjb->aicj", Ru, Rv)
    Th = torch.einsum("klij,aicj->akcl", G4, RR)  # [ru, su', rv, sv']
    Tm = Th.permute(3, 2, 1, 0).reshape(dv * rv, du * ru).T  # column-major (ru + r su') x (rv + r sv')
    U, s, Vh = torch.linalg.svd(Tm, full_matrices=False)
    keep, _ = O.truncation_keep((s.detach().numpy()) ** 2, cutoff, maxdim)
    U, s, Vh = U[:, :keep], s[:keep], Vh[:keep]
    sq = torch.sqrt(s)
    Fu = U * sq                       # rows (ru + r su'), columns j
    Fv = (sq[:, None] * Vh).T         # rows (rv + r sv'), columns j

    def rebuild(t, ax, Q, F, r, d, ext, Sinv):
        F3 = F.T.reshape(keep, d, r).permute(2, 1, 0)   # [r, s', j]
        core = torch.einsum("er,rsj->esj", Q, F3)       # [e, s', j], e column-major over ext
        shape = [t.shape[a] for a in ext]
        # e -> ext axes (column-major): reshape with reversed dims then reverse them
        core = core.permute(2, 1, 0).reshape(keep, d, *reversed(shape))
        core = core.permute(*reversed(range(core.dim())))  # [ext..., s', j]
        nd = t.dim()
        out = torch.movedim(core, list(range(nd)), ext + [0, ax])
        for a, si in Sinv.items():
            out = _mode(out, si.T, a)
        return out

    pu = rebuild(psi_u, ax_u, Qu, Fu, ru, du, ext_u, Sinv_u)
    pv = rebuild(psi_v, ax_v, Qv, Fv, rv, dv, ext_v, Sinv_v)
    signs = []
    for j in range(keep):
        sl = pu.detach().select(ax_u, j).permute(*reversed(range(pu.dim() - 1))).reshape(-1)
        signs.append(-1.0 if float(sl[torch.argmax(torch.abs(sl))]) < 0 else 1.0)
    sg = torch.tensor(signs, dtype=pu.dtype)
    shape_u = [1] * pu.dim()
    shape_u[ax_u] = keep
    shape_v = [1] * pv.dim()
    shape_v[ax_v] = keep
    pu = pu * sg.reshape(shape_u)
    pv = pv * sg.reshape(shape_v)
    if normalize:
        pu = pu / torch.linalg.norm(pu)
        pv = pv / torch.linalg.norm(pv)
    return pu, pv
