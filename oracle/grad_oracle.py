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
