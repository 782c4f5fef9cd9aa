"""The C++ caller of the C ABI (tests/cpp/abi_driver.cpp): it compiles and
links against include/itnbp.h and libitnbp.so with g++ (CPU), and on the GPU
its BP messages, <Sz>, logscalar and forest-cover order match the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import itnbp
from helpers import HEADER, LIB, ROOT, oracle_graph, oracle_psi
from oracle import bp_oracle as O

DRIVER = os.path.join(ROOT, "tests", "cpp", "abi_driver.cpp")


def build_driver(tmp):
    exe = os.path.join(tmp, "abi_driver")
    libdir = os.path.dirname(LIB)
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.dirname(HEADER), DRIVER, "-o", exe,
           "-L" + libdir, "-litnbp", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_driver_compiles_and_links(tmp_path):
    assert os.path.exists(build_driver(str(tmp_path)))


def write_input(path, g, psi):
    edges, inc_ptr, inc_edge = g.csr()
    chi = np.array([psi.bond_dim(a, b) for (a, b) in g.edges], dtype=np.int32)
    phys = np.array([psi.phys_dim(v) for v in g.vertices], dtype=np.int32)
    flat = np.concatenate([np.asarray(psi[v], dtype=np.float64).reshape(-1, order="F") for v in g.vertices])
    with open(path, "wb") as f:
        f.write(np.array([g.nv, g.ne], dtype=np.int32).tobytes())
        for a in (edges, inc_ptr, inc_edge, chi, phys):
            f.write(np.ascontiguousarray(a, dtype=np.int32).tobytes())
        f.write(flat.tobytes())


@pytest.mark.gpu
def test_driver_bp_matches_oracle(tmp_path):
    exe = build_driver(str(tmp_path))
    g = itnbp.named_grid((4, 4))
    psi = itnbp.random_tensornetwork(np.random.default_rng(1234), np.float64, g, link_space=2)
    inp = os.path.join(str(tmp_path), "in.bin")
    write_input(inp, g, psi)
    r = subprocess.run([exe, inp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    assert out["no_maxiter_rc"] == 1  # ITNBP_ERR_ARG: "You need to specify a number of iterations for BP!"
    og, opsi = oracle_graph(g), oracle_psi(psi)
    ref, it, _, _ = O.bp_update(og, opsi, O.initialize_messages(og, opsi), maxiter=30, tol=1e-12)
    assert out["iterations"] == it
    flat = np.array(out["messages"])
    for m in range(2 * g.ne):
        a, b = g.message_edge(m)
        got = flat[4 * m:4 * m + 4].reshape(2, 2, order="F")
        np.testing.assert_allclose(got, ref[(g.index[a], g.index[b])], atol=1e-10)
    for i in range(g.nv):
        assert out["sz"][i] == pytest.approx(O.expect1(O.SZ, O.rdm1(og, opsi, ref, i)), abs=1e-10)
    assert out["logscalar"][0] == pytest.approx(np.real(O.logscalar(og, opsi, ref)), rel=1e-10)
    seq = [g.message_edge(m) for m in out["sequence"]]
    assert [(g.index[a], g.index[b]) for (a, b) in seq] == O.forest_cover_sequence(og)
