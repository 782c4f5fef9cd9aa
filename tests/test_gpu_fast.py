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


@pytest.mark.parametrize("pipeline,lds", [("eig", "0"), ("legacy", "0"), ("legacy", "1")])
def test_f64_pipeline_matches_oracle(net34, pipeline, lds, monkeypatch):
    monkeypatch.setenv("ITNBP_F64_LDS", lds)
    monkeypatch.setenv("ITNBP_FAST_LEGACY", "1" if pipeline == "legacy" else "0")
    g, psi = net34
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=3, edge_sequence="parallel")
    assert "mfma-f64" in bpc.last_update["path"]
    assert ("eig" in bpc.last_update["path"]) == (pipeline == "eig")
    og, opsi = oracle_graph(g), oracle_psi(psi)
    ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=3, schedule="parallel")
    got = to_index(g, bpc.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    assert err < 1e-12, err


@pytest.mark.parametrize("flavour,tol_low", [("eig", 1e-3), ("eig_bf16", 3e-3), ("eig_x3", 1e-7),
                                             ("precise", 2e-4), ("compact", 3e-3)])
def test_bf16_pipeline_then_f64_refinement(net34, oracle_ref, flavour, tol_low, monkeypatch):
    """BF16 site tensors; eig: FP16 intermediates (default), eig_bf16: BF16 ones,
    eig_x3: the compensated flavour (precision "bf16x3", operands split into
    three BF16 terms, FP32 intermediates); precise / compact: the absorbed-pair passes."""
    monkeypatch.setenv("ITNBP_BF16_PRECISE", "1" if flavour == "precise" else "0")
    monkeypatch.setenv("ITNBP_FAST_LEGACY", "0" if flavour.startswith("eig") else "1")
    monkeypatch.setenv("ITNBP_EIG_BF16", "1" if flavour == "eig_bf16" else "0")
    low_prec, full_prec = ("bf16x3", "bf16x3_f64") if flavour == "eig_x3" else ("f32", "bf16_f64")
    flavour = "eig" if flavour.startswith("eig") else flavour
    g, psi = net34
    m3, m40 = oracle_ref
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    low = itnbp.update(bpc, maxiter=3, edge_sequence="parallel", precision=low_prec)
    assert "bf16" in low.last_update["path"]
    assert ("eig" in low.last_update["path"]) == (flavour == "eig")
    assert ("bf16x3" in low.last_update["path"]) == (low_prec == "bf16x3")
    assert ("bf16x2" in low.last_update["path"]) == (flavour == "precise")
    got = to_index(g, low.messages())
    err_low = max(np.abs(got[e] - m3[e]).max() for e in m3)
    print(f"{flavour}: max message error after 3 low-precision sweeps {err_low:.3g}")
    assert err_low < tol_low, err_low
    full = itnbp.update(bpc, maxiter=40, res_tol=1e-11, edge_sequence="parallel", precision=full_prec,
                        switch_tol=1e-4)
    assert full.last_update["iterations_low"] >= 1 and full.last_update["converged"]
    got = to_index(g, full.messages())
    err = max(np.abs(got[e] - m40[e]).max() for e in m40)
    assert err < 1e-10, err


@pytest.mark.parametrize("scale", [2.0 ** -40, 1e-9, 3e5])
def test_eig16_site_scales(net34, scale):
    """The 16-bit eigenbasis pipeline multiplies each site by a power of two that
    brings its largest entry into [1/2, 1) (the FP16 intermediates' range) and
    divides the Grams by its square: normalised messages of psi * scale equal
    those of psi (bit for bit for a power of two), and unnormalised ones carry
    scale^2 (one Jacobi sweep against the FP64 pipeline)."""
    g, psi = net34
    base = itnbp.update(itnbp.initialize_cache(psi, storage="bf16"), maxiter=2, edge_sequence="parallel",
                        precision="f32")
    assert "bf16-eig" in base.last_update["path"]
    psi_s = psi.copy()
    for v in g.vertices:
        psi_s[v] = psi[v] * scale
    got = itnbp.update(itnbp.initialize_cache(psi_s, storage="bf16"), maxiter=2, edge_sequence="parallel",
                       precision="f32")
    a, b = base.messages(), got.messages()
    err = max(np.abs(np.asarray(a[e]) - np.asarray(b[e])).max() for e in a)
    print(f"scale {scale:g}: normalised messages differ by {err:.3g}")
    if np.log2(scale) == np.round(np.log2(scale)):
        assert err == 0.0, err
    else:
        assert err < 5e-3, err
    raw = dict(message_update_alg={"alg": "contract", "normalize": False})
    bpc = itnbp.initialize_cache(psi_s, storage="bf16")
    lo = itnbp.update(bpc, maxiter=1, edge_sequence="parallel", precision="f32", **raw)
    hi = itnbp.update(bpc, maxiter=1, edge_sequence="parallel", precision="f64", **raw)
    assert "bf16-eig" in lo.last_update["path"] and "f64-eig" in hi.last_update["path"]
    a, b = lo.messages(), hi.messages()
    rel = max(np.abs(np.asarray(a[e], dtype=np.float64) - np.asarray(b[e])).max() / np.abs(b[e]).max() for e in b)
    print(f"scale {scale:g}: unnormalised messages, relative difference {rel:.3g}")
    assert rel < 1e-2, rel


def _random_messages(g, psi, rng, symmetric):
    """Random (indefinite when symmetric) messages for every directed edge."""
    msgs = {}
    for (u, v) in g.edges:
        chi = psi[u].shape[g.axis(u, v)]
        for e in ((u, v), (v, u)):
            a = rng.standard_normal((chi, chi))
            msgs[e] = (a + a.T) / 2 if symmetric else a
    return msgs


@pytest.mark.parametrize("symmetric", [True, False])
def test_user_messages_pick_pipeline_and_match_oracle(net34, symmetric):
    """Symmetric user messages (eigenvalues of both signs) run on the eigenbasis
    pipeline; a non-symmetric one sends the sweep to the absorbed-pair passes.
    Both match the oracle started from the same messages."""
    g, psi = net34
    rng = np.random.default_rng(7 if symmetric else 8)
    msgs = _random_messages(g, psi, rng, symmetric)
    bpc = itnbp.initialize_cache(psi)
    for e, m in msgs.items():
        bpc.set_message_(e, m)
    out = itnbp.update(bpc, maxiter=2, edge_sequence="parallel")
    path = out.last_update["path"]
    assert "mfma-f64" in path
    assert ("eig" in path) == symmetric
    og, opsi = oracle_graph(g), oracle_psi(psi)
    start = {(g.index[a], g.index[b]): m for (a, b), m in msgs.items()}
    ref, _, _, _ = O.bp_update(og, opsi, start, maxiter=2, schedule="parallel")
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    assert err < 1e-12, err


def test_progressive_precision_reaches_the_bf16x3_floor(net34, oracle_ref):
    """Precision "bf16_bf16x3": one-term FP16 sweeps until the residual < 1e-2,
    then three-term sweeps; after 40 sweeps the messages are within the
    three-term flavour's bound of the oracle's fixed point."""
    g, psi = net34
    m3, m40 = oracle_ref
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    out = itnbp.update(bpc, maxiter=40, res_tol=1e-9, edge_sequence="parallel", precision="bf16_bf16x3",
                       switch_tol=1e-2)
    path, lu = out.last_update["path"], out.last_update
    assert path.startswith("mfma-bf16-eig-chi32+mfma-bf16x3-eig-chi32"), path
    assert lu["iterations_low"] == lu["iterations"] >= 3
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - m40[e]).max() for e in m40)
    print(f"progressive: {lu['iterations']} sweeps, residual {lu['max_residual']:.3g}, max message error {err:.3g}")
    assert err < 1e-6, err


def test_three_phase_precision_refines_to_fp64(net34, oracle_ref):
    """Precision "bf16_bf16x3_f64": FP16 sweeps to residual 1e-2, three-term sweeps
    to switch_tol, then FP64 to 1e-11; the messages match the oracle's fixed point
    to 1e-10."""
    g, psi = net34
    m3, m40 = oracle_ref
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    out = itnbp.update(bpc, maxiter=40, res_tol=1e-11, edge_sequence="parallel", precision="bf16_bf16x3_f64",
                       switch_tol=1e-6)
    path, lu = out.last_update["path"], out.last_update
    assert path.startswith("mfma-bf16-eig-chi32+mfma-bf16x3-eig-chi32+mfma-f64-eig-chi32"), path
    assert lu["converged"] and 2 <= lu["iterations_low"] < lu["iterations"]
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - m40[e]).max() for e in m40)
    print(f"three-phase: {lu['iterations_low']} + {lu['iterations'] - lu['iterations_low']} sweeps, error {err:.3g}")
    assert err < 1e-10, err


@pytest.mark.parametrize("symmetric", [True, False])
def test_bf16x3_follows_message_symmetry(net34, symmetric):
    """Precision "bf16x3" with user messages: symmetric ones run the three-term
    eigenbasis kernels, a non-symmetric one the absorbed-pair passes with two-term
    splits and FP32 intermediates. After two Jacobi sweeps on the same BF16 tensors
    from random messages they are within 1e-6 (three terms, measured 1.2e-7) and
    5e-6 (two terms, measured 1.2e-6) of the oracle."""
    g, psi = net34
    rng = np.random.default_rng(17 if symmetric else 18)
    msgs = _random_messages(g, psi, rng, symmetric)
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    for e, m in msgs.items():
        bpc.set_message_(e, m)
    out = itnbp.update(bpc, maxiter=2, edge_sequence="parallel", precision="bf16x3")
    path = out.last_update["path"]
    assert path.startswith("mfma-bf16x3-eig-chi32" if symmetric else "mfma-bf16x2-chi32"), path
    og, opsi = oracle_graph(g), oracle_psi(bpc.tensornetwork())
    start = {(g.index[a], g.index[b]): m for (a, b), m in msgs.items()}
    ref, _, _, _ = O.bp_update(og, opsi, start, maxiter=2, schedule="parallel")
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    print(f"symmetric={symmetric}: max message error {err:.3g}")
    assert err < (1e-6 if symmetric else 5e-6), err


@pytest.mark.parametrize("precision", ["f64", "bf16_f64", "bf16x3_f64"])
def test_degree3_only_lattice(precision):
    """Heavy-hex lattice with chi = 32: degree-3 and degree-2 sites only, so the
    eigenbasis pipeline runs its degree-3 kernels without any bulk vertex."""
    g = itnbp.heavy_hex(4, 3)
    assert max(g.degree(v) for v in g.vertices) == 3
    psi = itnbp.random_tensornetwork(np.random.default_rng(33), np.float64, g, link_space=32,
                                     scale=1 / np.sqrt(32))
    storage = "f64" if precision == "f64" else "bf16"
    bpc = itnbp.initialize_cache(psi, storage=storage)
    psi_b = bpc.tensornetwork()
    og, opsi = oracle_graph(g), oracle_psi(psi_b)
    if precision == "f64":
        out = itnbp.update(bpc, maxiter=3, edge_sequence="parallel")
        assert "mfma-f64-eig" in out.last_update["path"]
        ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=3, schedule="parallel")
        tol = 1e-12
    else:
        x3 = precision == "bf16x3_f64"
        low = itnbp.update(bpc, maxiter=3, edge_sequence="parallel", precision="bf16x3" if x3 else "f32")
        assert ("bf16x3-eig" if x3 else "bf16-eig") in low.last_update["path"]
        ref3, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=3, schedule="parallel")
        got = to_index(g, low.messages())
        err_low = max(np.abs(got[e] - ref3[e]).max() for e in ref3)
        print(f"{precision}: max message error after 3 low-precision sweeps {err_low:.3g}")
        assert err_low < (1e-7 if x3 else 3e-3), err_low
        out = itnbp.update(bpc, maxiter=60, res_tol=1e-11, edge_sequence="parallel", precision=precision,
                           switch_tol=1e-4)
        assert out.last_update["converged"]
        it = out.last_update["iterations"]
        ref, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=80, schedule="parallel")
        tol = 1e-9
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    assert err < tol, err


def test_site_rdm_eigenbasis_matches_oracle_and_generic(net34, monkeypatch):
    """One-site density matrices, <X>, <Z>, vertex scalars and logscalar of a chi = 32
    3x4 grid (two bulk, six degree-3 sites) through the rotated-Gram route
    (rho = sum_r prod lam Y[s, r] Y[s', r] in the eigenbasis of every incoming
    message) against the oracle's open contraction and the generic kernels."""
    g, psi = net34
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=4, edge_sequence="parallel")
    og, opsi = oracle_graph(g), oracle_psi(psi)
    ref = to_index(g, bpc.messages())
    ctx = bpc.ctx
    monkeypatch.setenv("ITNBP_RDM_FAST_MIN", "1")
    ctx.profile(True)
    for v in g.vertices:
        r = itnbp.rdm(bpc, [v])
        rr = O.rdm1(og, opsi, ref, g.index[v])
        np.testing.assert_allclose(r, rr, rtol=1e-11, atol=1e-12 * np.abs(rr).max())
    vs, _ = itnbp.region_scalars(bpc)
    n_fast, _ = ctx.kernel_time("rdm1_eig")
    ctx.profile(False)
    assert n_fast == len(g.vertices) - 4 + 1  # every bulk and degree-3 site, then the scalars
    for i in range(g.nv):
        assert vs[i] == pytest.approx(O.vertex_scalar(og, opsi, ref, i), rel=1e-11)
    assert itnbp.logscalar(bpc) == pytest.approx(O.logscalar(og, opsi, ref), rel=1e-11)
    fast = {op: itnbp.expect_on_cache(bpc, op) for op in ("X", "Z")}
    monkeypatch.setenv("ITNBP_DISABLE_FAST", "1")
    for op in ("X", "Z"):
        gen = itnbp.expect_on_cache(bpc, op)
        err = max(abs(fast[op][v] - gen[v]) for v in g.vertices)
        assert err < 1e-12, (op, err)


def test_pair_rdm_eigenbasis_matches_oracle_and_generic(net34, monkeypatch):
    """Two-site density matrices and <ZZ> of a chi = 32 3x4 grid through the
    eigenbasis half-pairs (a 64 x 64 Gram of the rotated site over every index but
    (s, open bond), rotated back by that bond's eigenvectors; corners stay on the
    generic kernels) against the oracle and the generic kernels."""
    g, psi = net34
    bpc = itnbp.update(itnbp.initialize_cache(psi), maxiter=4, edge_sequence="parallel")
    og, opsi = oracle_graph(g), oracle_psi(psi)
    ref = to_index(g, bpc.messages())
    ctx = bpc.ctx
    monkeypatch.setenv("ITNBP_RDM_FAST_MIN", "1")
    ctx.profile(True)
    for (a, b) in g.edges:
        r = itnbp.rdm(bpc, [a, b])
        rr = O.rdm2(og, opsi, ref, g.index[a], g.index[b])
        np.testing.assert_allclose(r, rr, rtol=1e-10, atol=1e-11 * np.abs(rr).max())
    n_fast, _ = ctx.kernel_time("rdm2_eig")
    ctx.profile(False)
    assert n_fast == len(g.edges)
    zz = O.two_site(O.Z, O.Z)
    fast = itnbp.expect2_on_cache(bpc, zz)
    monkeypatch.setenv("ITNBP_DISABLE_FAST", "1")
    gen = itnbp.expect2_on_cache(bpc, zz)
    err = max(abs(fast[e] - gen[e]) for e in gen)
    assert err < 1e-12, err


def test_density_matrix_routes_from_bf16_storage(net34, monkeypatch):
    """The chi = 32 one- and two-site routes build their FP64 layouts from BF16 site
    storage: <X>, <ZZ> and the vertex scalars agree with the generic kernels on the
    same (BF16-rounded) tensors."""
    g, psi = net34
    bpc = itnbp.update(itnbp.initialize_cache(psi, storage="bf16"), maxiter=4, edge_sequence="parallel",
                       precision="f64")
    zz = O.two_site(O.Z, O.Z)
    monkeypatch.setenv("ITNBP_RDM_FAST_MIN", "1")
    bpc.ctx.profile(True)
    fx = itnbp.expect_on_cache(bpc, "X")
    fzz = itnbp.expect2_on_cache(bpc, zz)
    fvs, _ = itnbp.region_scalars(bpc)
    n1, _ = bpc.ctx.kernel_time("rdm1_eig")
    n2, _ = bpc.ctx.kernel_time("rdm2_eig")
    bpc.ctx.profile(False)
    assert n1 == 2 and n2 == 1  # <X> and the scalars; <ZZ>
    monkeypatch.setenv("ITNBP_DISABLE_FAST", "1")
    gx = itnbp.expect_on_cache(bpc, "X")
    gzz = itnbp.expect2_on_cache(bpc, zz)
    gvs, _ = itnbp.region_scalars(bpc)
    assert max(abs(fx[v] - gx[v]) for v in gx) < 1e-12
    assert max(abs(fzz[e] - gzz[e]) for e in gzz) < 1e-12
    np.testing.assert_allclose(fvs, gvs, rtol=1e-11)


@pytest.mark.parametrize("symmetric", [True, False])
def test_f16x2_flavour_matches_oracle(net34, symmetric):
    """Precision "f16x2": symmetric messages run the eigenbasis kernels with every
    rounded operand split into two FP16 terms (FP32 intermediates); a non-symmetric
    message takes the absorbed-pair passes with two-term BF16 splits.  Two Jacobi
    sweeps from random messages against the oracle on the same BF16 tensors:
    within 1e-6 (a CPU rounding model of the two-FP16-term pipeline gives 2e-7 to
    4e-7 per sweep) and 5e-6."""
    g, psi = net34
    rng = np.random.default_rng(27 if symmetric else 28)
    msgs = _random_messages(g, psi, rng, symmetric)
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    for e, m in msgs.items():
        bpc.set_message_(e, m)
    out = itnbp.update(bpc, maxiter=2, edge_sequence="parallel", precision="f16x2")
    path = out.last_update["path"]
    assert path.startswith("mfma-f16x2-eig-chi32" if symmetric else "mfma-bf16x2-chi32"), path
    og, opsi = oracle_graph(g), oracle_psi(bpc.tensornetwork())
    start = {(g.index[a], g.index[b]): m for (a, b), m in msgs.items()}
    ref, _, _, _ = O.bp_update(og, opsi, start, maxiter=2, schedule="parallel")
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - ref[e]).max() for e in ref)
    print(f"f16x2 symmetric={symmetric}: max message error {err:.3g}")
    assert err < (1e-6 if symmetric else 5e-6), err


def test_progressive_f16x2_reaches_its_floor(net34, oracle_ref):
    """Precision "bf16_f16x2": one-term FP16 sweeps until the residual < 1e-2, then
    two-FP16-term sweeps; after 40 sweeps within 2e-6 of the oracle's fixed point."""
    g, psi = net34
    m3, m40 = oracle_ref
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    out = itnbp.update(bpc, maxiter=40, res_tol=1e-9, edge_sequence="parallel", precision="bf16_f16x2",
                       switch_tol=1e-2)
    path, lu = out.last_update["path"], out.last_update
    assert path.startswith("mfma-bf16-eig-chi32+mfma-f16x2-eig-chi32"), path
    got = to_index(g, out.messages())
    err = max(np.abs(got[e] - m40[e]).max() for e in m40)
    print(f"bf16_f16x2: {lu['iterations']} sweeps, residual {lu['max_residual']:.3g}, max message error {err:.3g}")
    assert err < 2e-6, err


def test_f16x2_degree3_only_lattice():
    """Heavy-hex, chi = 32: the degree-3 kernels of the two-FP16-term flavour."""
    g = itnbp.heavy_hex(4, 3)
    psi = itnbp.random_tensornetwork(np.random.default_rng(33), np.float64, g, link_space=32,
                                     scale=1 / np.sqrt(32))
    bpc = itnbp.initialize_cache(psi, storage="bf16")
    og, opsi = oracle_graph(g), oracle_psi(bpc.tensornetwork())
    low = itnbp.update(bpc, maxiter=3, edge_sequence="parallel", precision="f16x2")
    assert "f16x2-eig" in low.last_update["path"], low.last_update["path"]
    ref3, _, _, _ = O.bp_update(og, opsi, O.identity_messages(og, opsi), maxiter=3, schedule="parallel")
    got = to_index(g, low.messages())
    err = max(np.abs(got[e] - ref3[e]).max() for e in ref3)
    print(f"f16x2 degree-3 lattice: max message error after 3 sweeps {err:.3g}")
    assert err < 1e-6, err
