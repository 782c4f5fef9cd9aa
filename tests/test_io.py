"""Interchange file format: round trips of graphs, site tensors (FP64, FP32,
BF16) and messages, byte-stable re-saves, and rejection of foreign files."""
import numpy as np
import pytest

import itnbp
from itnbp.cache import bf16_round


def _same_graph(g1, g2):
    assert g1.vertices == g2.vertices
    assert g1.edges == g2.edges
    for v in g1.vertices:
        assert g1.incident(v) == g2.incident(v)


@pytest.mark.parametrize("make,dtype", [
    (lambda: itnbp.named_grid((3, 4)), np.float64),
    (lambda: itnbp.heavy_hex(4, 3), np.float32),
    (lambda: itnbp.comb_tree(3, 3), np.float64),
])
def test_round_trip_tensors_and_messages(tmp_path, make, dtype):
    g = make()
    rng = np.random.default_rng(0)
    chis = [int(c) for c in rng.integers(1, 5, size=g.ne)]
    psi = itnbp.random_tensornetwork(rng, dtype, g, link_space=chis)
    msgs = {}
    for (a, b) in g.edges:
        chi = psi[a].shape[g.axis(a, b)]
        msgs[(a, b)] = rng.standard_normal((chi, chi))
        msgs[(b, a)] = rng.standard_normal((chi, chi))
    path = tmp_path / "net.itnbp"
    itnbp.save_network(path, psi, msgs)
    psi2, msgs2 = itnbp.load_network(path)
    _same_graph(g, psi2.graph)
    for v in g.vertices:
        assert psi2[v].dtype == np.dtype(dtype)
        np.testing.assert_array_equal(psi2[v], psi[v])
    assert set(msgs2) == set(msgs)
    for e in msgs:
        np.testing.assert_array_equal(msgs2[e], msgs[e])
    # a re-save of what was loaded is byte-identical
    path2 = tmp_path / "net2.itnbp"
    itnbp.save_network(path2, psi2, msgs2)
    assert path.read_bytes() == path2.read_bytes()


def test_bf16_storage_and_no_messages(tmp_path):
    g = itnbp.named_grid((2, 3))
    psi = itnbp.random_tensornetwork(np.random.default_rng(1), np.float64, g, link_space=3)
    path = tmp_path / "b.itnbp"
    itnbp.save_network(path, psi, bf16=True)
    psi2, msgs2 = itnbp.load_network(path)
    assert msgs2 is None
    for v in g.vertices:
        ref = bf16_round(psi[v].reshape(-1, order="F"))[0].astype(np.float64).reshape(psi[v].shape, order="F")
        np.testing.assert_array_equal(psi2[v], ref)


def test_layout_is_column_major_with_json_header(tmp_path):
    """The payload is readable without the package: header offsets point at
    column-major [s, b_1 .. b_k] little-endian data."""
    import json
    import struct
    g = itnbp.path_graph(3)
    psi = itnbp.random_tensornetwork(np.random.default_rng(2), np.float64, g, link_space=2)
    path = tmp_path / "p.itnbp"
    itnbp.save_network(path, psi)
    raw = path.read_bytes()
    assert raw[:8] == b"ITNBPTN1"
    (hl,) = struct.unpack("<Q", raw[8:16])
    h = json.loads(raw[16:16 + hl])
    v = g.vertices[1]
    t = h["tensors"][1]
    arr = np.frombuffer(raw, dtype="<f8", count=int(np.prod(t["shape"])), offset=16 + hl + t["offset"])
    np.testing.assert_array_equal(arr, psi[v].reshape(-1, order="F"))
    assert t["offset"] % 64 == 0


def test_foreign_file_rejected(tmp_path):
    path = tmp_path / "x.itnbp"
    path.write_bytes(b"NOTAFILE" + b"\0" * 32)
    with pytest.raises(ValueError):
        itnbp.load_network(path)
