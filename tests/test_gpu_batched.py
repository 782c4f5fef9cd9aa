"""GPU parity of batched heterogeneous BP (several lattices and bond dimensions
in one sweep launch) against the CPU oracle run per instance."""
import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, oracle_psi, to_index
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


def instances(dtype, seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for g, chi in [(itnbp.named_grid((3, 3)), 2), (itnbp.named_grid((4, 3)), 3), (itnbp.heavy_hex(3, 2), 4),
                   (itnbp.path_graph(5), 3), (itnbp.named_grid((2, 5)), 6), (itnbp.named_grid((3, 3)), 16)]:
        psi = itnbp.random_tensornetwork(rng, dtype, g, link_space=chi, scale=1 / np.sqrt(chi))
        out.append((g, psi))
    return out


@pytest.mark.parametrize("dtype,prec,tol", [(np.float64, "f64", 1e-11), (np.float32, "f32", 2e-5)])
def test_batched_matches_per_instance_oracle(dtype, prec, tol):
    inst = instances(dtype)
    caches = [itnbp.BeliefPropagationCache(psi) for _, psi in inst]
    for c, (g, psi) in zip(caches, inst):
        c.set_messages_(itnbp.identity_messages(psi))
    res = itnbp.update_batched_(caches, maxiter=25, res_tol=1e-9 if prec == "f64" else 1e-5, precision=prec)
    for c, r, (g, psi) in zip(caches, res, inst):
        og, opsi = oracle_graph(g), oracle_psi(psi)
        ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=r["iterations"],
                                   schedule="parallel")
        got = to_index(g, c.messages())
        assert max(np.abs(got[e] - ref[e]).max() for e in ref) < tol
    # instances converge at different sweeps and drop out
    assert len({r["iterations"] for r in res}) > 1
    # small-chi sites on the per-vertex kernel, the chi = 16 degree-4 centre on the generic one
    assert "[vertex " in res[0]["path"], res[0]["path"]


def test_batched_with_a_non_symmetric_message_uses_generic_kernels():
    """One non-symmetric user message in one instance keeps every instance of the
    batch on the generic kernels (the per-vertex kernel needs symmetric messages);
    the results still match the oracle."""
    inst = instances(np.float64, seed=4)[:3]
    caches = [itnbp.BeliefPropagationCache(psi) for _, psi in inst]
    starts = []
    rng = np.random.default_rng(9)
    for k, (c, (g, psi)) in enumerate(zip(caches, inst)):
        msgs = itnbp.identity_messages(psi)
        if k == 1:
            e = next(iter(msgs))
            msgs[e] = msgs[e] + 0.1 * rng.standard_normal(msgs[e].shape)
        c.set_messages_(msgs)
        starts.append({(g.index[a], g.index[b]): np.asarray(m) for (a, b), m in msgs.items()})
    res = itnbp.update_batched_(caches, maxiter=6, precision="f64")
    assert "[vertex" not in res[0]["path"], res[0]["path"]
    for c, r, (g, psi), start in zip(caches, res, inst, starts):
        og, opsi = oracle_graph(g), oracle_psi(psi)
        ref, _, _, _ = O.bp_update(og, opsi, start, maxiter=r["iterations"], schedule="parallel")
        got = to_index(g, c.messages())
        assert max(np.abs(got[e] - ref[e]).max() for e in ref) < 1e-11


def run_batched(inst, monkeypatch, graph, **kw):
    monkeypatch.setenv("ITNBP_BATCH_GRAPH", "1" if graph else "0")
    caches = [itnbp.BeliefPropagationCache(psi) for _, psi in inst]
    for c, (g, psi) in zip(caches, inst):
        c.set_messages_(itnbp.identity_messages(psi))
    res = itnbp.update_batched_(caches, **kw)
    return res, [to_index(g, c.messages()) for c, (g, _) in zip(caches, inst)]


@pytest.mark.parametrize("kw", [dict(maxiter=25, res_tol=1e-9, precision="f64"),
                                dict(maxiter=13, precision="f64"),
                                dict(maxiter=30, tol=1e-10, precision="f32")])
def test_graph_loop_matches_the_host_loop(monkeypatch, kw):
    """The device-resident loop (eight sweeps captured in a HIP graph and
    replayed, per-instance convergence and drop-out decided on the device, the
    large-site Gram tile autotuned) stops every instance at the sweep the
    host-driven loop does, with the same messages; without a tolerance every
    instance runs maxiter sweeps, also when maxiter is not a multiple of eight."""
    inst = instances(np.float32 if kw["precision"] == "f32" else np.float64)
    rg, mg = run_batched(inst, monkeypatch, True, **kw)
    rh, mh = run_batched(inst, monkeypatch, False, **kw)
    assert rg[0]["path"].startswith("batched-graph [vertex "), rg[0]["path"]
    assert rh[0]["path"].startswith("batched-generic"), rh[0]["path"]
    assert [r["iterations"] for r in rg] == [r["iterations"] for r in rh]
    assert [r["converged"] for r in rg] == [r["converged"] for r in rh]
    if "tol" not in kw and "res_tol" not in kw:
        assert all(r["iterations"] == kw["maxiter"] for r in rg)
    f64 = kw["precision"] == "f64"
    atol = 1e-11 if f64 else 1e-5
    for a, b in zip(mg, mh):
        assert max(np.abs(a[e] - b[e]).max() for e in a) < atol
    for r_g, r_h in zip(rg, rh):
        assert r_g["mean_diff"] == pytest.approx(r_h["mean_diff"], rel=1e-6 if f64 else 1e-2,
                                                 abs=1e-14 if f64 else 1e-7)


@pytest.mark.parametrize("tile", ["1", "2", "4", "192"])
def test_large_site_kernel_variants_match_the_oracle(monkeypatch, tile):
    """Every variant of the large-site kernel (k_vertex_big: sites whose tensor
    does not fit in LDS), forced through ITNBP_VERTEX_BIG_TILE: the FP32 MFMA
    kernel with 1, 2 or 4 tiles in flight per wave and the VALU kernel with a
    192-fibre Gram tile, on the chi = 16 degree-4 centre of a 3x3 grid and the
    four chi = 12 degree-4 sites of a 4x4 grid, FP32, against the oracle."""
    monkeypatch.setenv("ITNBP_VERTEX_BIG_TILE", tile)
    rng = np.random.default_rng(17)
    inst = []
    for g, chi in [(itnbp.named_grid((3, 3)), 16), (itnbp.named_grid((4, 4)), 12)]:
        inst.append((g, itnbp.random_tensornetwork(rng, np.float32, g, link_space=chi, scale=1 / np.sqrt(chi))))
    res, got = run_batched(inst, monkeypatch, True, maxiter=12, precision="f32")
    assert "[vertex " in res[0]["path"], res[0]["path"]
    for r, m, (g, psi) in zip(res, got, inst):
        og, opsi = oracle_graph(g), oracle_psi(psi)
        ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=r["iterations"],
                                   schedule="parallel")
        err = max(np.abs(m[e] - ref[e]).max() for e in ref)
        print(f"tile {tile}: {g.nv} sites, {r['iterations']} sweeps, max abs err {err:.2e}, path {r['path']}",
              flush=True)
        assert err < 2e-5


@pytest.mark.parametrize("queue", ["0", "1"])
@pytest.mark.parametrize("graph", [True, False])
def test_persistent_small_site_launch_matches_the_oracle(monkeypatch, queue, graph):
    """The LDS-resident sites as one workgroup per task and as the persistent
    queue kernel (a resident grid taking the cost-sorted tasks from a device
    counter), forced through ITNBP_VERTEX_QUEUE, in the graph loop and the host
    loop, FP64 against the oracle."""
    monkeypatch.setenv("ITNBP_VERTEX_QUEUE", queue)
    inst = instances(np.float64, seed=5)
    res, got = run_batched(inst, monkeypatch, graph, maxiter=15, precision="f64")
    assert "[vertex " in res[0]["path"], res[0]["path"]
    if graph:
        assert ("persistent" in res[0]["path"]) == (queue == "1"), res[0]["path"]
    for r, m, (g, psi) in zip(res, got, inst):
        og, opsi = oracle_graph(g), oracle_psi(psi)
        ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=r["iterations"],
                                   schedule="parallel")
        assert max(np.abs(m[e] - ref[e]).max() for e in ref) < 1e-11
