"""GPU parity of the chi=32 bulk-vertex MFMA pipeline (FP64 and both BF16
flavours) against the CPU oracle."""
import os

import numpy as np
import pytest

import itnbp
from helpers import oracle_graph, oracle_psi, to_index
from oracle import bp_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net34():
    g = itnbp.named_grid((3, 4))  # two bulk (degree-4) vertices
    psi = itnbp.random_tensornetwork(np.random.default_rng(32), np.float64, g, link_space=32,
                                     scale=1 / np.sqrt(32))
    return g, psi


@pytest.fixture(scope="module")
def oracle_ref(net34):
    g, psi = net34
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    psi_b = bpc.tensornetwork()
    og, opsi = oracle_graph(g), oracle_psi(psi_b)
    m3, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=3, schedule="parallel")
    m40, _, _, _ = O.bp_update(og, opsi, m3, maxiter=37, schedule="parallel")
    return m3, m40


@pytest.mark.parametrize("lds", ["0", "1"])
def test_f64_pipeline_matches_oracle(net34, lds, monkeypatch):
    monkeypatch.setenv("ITNBP_F64_LDS", lds)
    g, psi = net34
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=3, edge_sequence="parallel")
    assert "mfma-f64" in bpc.last_update["path"]
    og, opsi = oracle_graph(g), oracle_psi(psi)
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=3, schedule="parallel")
    got = to_index(g, bpc.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    assert err < 1e-12, err


@pytest.mark.parametrize("flavour,tol_low", [("precise", 2e-4), ("compact", 3e-3)])
def test_bf16_pipeline_then_f64_refinement(net34, oracle_ref, flavour, tol_low, monkeypatch):
    monkeypatch.setenv("ITNBP_BF16_PRECISE", "1" if flavour == "precise" else "0")
    g, psi = net34
    m3, m40 = oracle_ref
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    low = itnbp.update(bpc, maxiter=3, edge_sequence="parallel", precision="f32")
    assert "bf16" in low.last_update["path"]
    got = to_index(g, low.messages())
    err_low = max(np.abs(got[e] - m3[e]).max() for e in m3)
    assert err_low < tol_low, err_low
    full = itnbp.update(bpc, maxiter=40, res_tol=1e-11, edge_sequence="parallel", precision="bf16_f64",
                        switch_tol=1e-4)
    assert full.last_update["iterations_low"] >= 1 and full.last_update["converged"]
    got = to_index(g, full.messages())
    err = max(np.abs(got[e] - m40[e]).max() for e in m40)
    assert err < 1e-10, err
