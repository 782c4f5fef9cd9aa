"""Interchange format for a tensor network state and its BP messages.

One file: an 8-byte magic ``ITNBPTN1``, the byte length of a UTF-8 JSON header
as a little-endian uint64, the header, then the raw little-endian payload
(every array 64-byte aligned, column-major).  The header carries everything a
reader needs without this package (a Julia or C reader is a few lines):

* ``vertices``: vertex names (tuples become JSON lists), in vertex order;
* ``edges``: ``[i, j]`` vertex indices, in edge order (message id 2e is i -> j,
  2e + 1 is j -> i, the C ABI's convention);
* ``incident``: per vertex, the edge ids of its bond axes in axis order;
* ``tensors``: per vertex ``{"offset", "shape"}``; axes ``[s, b_1 .. b_k]``, the
  site tensor layout of random_tensornetwork (test/utils.jl:23-38);
* ``messages`` (optional): per directed edge ``{"src", "dst", "offset", "chi"}``,
  ``M[b_ket, b_bra]`` in FP64.

BF16 site tensors are stored as their 16-bit patterns (``"dtype": "bfloat16"``).
"""
from __future__ import annotations

import json
import struct

import numpy as np

from .network import ITensorNetwork, NamedGraph

MAGIC = b"ITNBPTN1"
_ALIGN = 64


def _name_to_json(v):
    return list(v) if isinstance(v, tuple) else v


def _name_from_json(v):
    return tuple(_name_from_json(x) for x in v) if isinstance(v, list) else v


def save(path, psi: ITensorNetwork, messages=None, *, bf16=False):
    """Write ``psi`` (and optionally ``messages``, a dict {(u, v): chi x chi}) to
    ``path``.  ``bf16=True`` stores the site tensors rounded to bfloat16."""
    from .cache import bf16_round
    g = psi.graph
    vidx = g.index
    payload = bytearray()

    def put(arr):
        pad = (-len(payload)) % _ALIGN
        payload.extend(b"\0" * pad)
        off = len(payload)
        payload.extend(np.ascontiguousarray(arr).tobytes())
        return off

    dtype = "bfloat16" if bf16 else ("float32" if psi.dtype == np.float32 else "float64")
    tensors = []
    for v in g.vertices:
        t = np.asarray(psi[v])
        flat = t.reshape(-1, order="F")
        if bf16:
            data = bf16_round(flat.astype(np.float32))[1]
        elif dtype == "float32":
            data = flat.astype("<f4")
        else:
            data = flat.astype("<f8")
        tensors.append({"offset": put(data), "shape": list(t.shape)})
    header = {
        "format": "itnbp-tensornetwork",
        "version": 1,
        "dtype": dtype,
        "order": "column-major",
        "axes": "[s, b_1 .. b_k], bonds in incident order",
        "vertices": [_name_to_json(v) for v in g.vertices],
        "edges": [[vidx[a], vidx[b]] for (a, b) in g.edges],
        "incident": [[g.edge_id(v, w) for w in g.incident(v)] for v in g.vertices],
        "tensors": tensors,
        "messages": None,
    }
    if messages is not None:
        entries = []
        for (a, b), m in sorted(messages.items(), key=lambda kv: (vidx[kv[0][0]], vidx[kv[0][1]])):
            m = np.asarray(m, dtype=np.float64)
            entries.append({"src": vidx[a], "dst": vidx[b], "chi": int(m.shape[0]),
                            "offset": put(m.reshape(-1, order="F").astype("<f8"))})
        header["messages"] = {"layout": "M[b_ket, b_bra] column-major", "dtype": "float64", "entries": entries}
    hb = json.dumps(header, separators=(",", ":")).encode()
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(hb)))
        f.write(hb)
        f.write(bytes(payload))


def load(path):
    """Read a file written by :func:`save`: returns ``(psi, messages)`` with
    ``messages`` None when the file holds none.  BF16 tensors are returned as
    their exact float64 values."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != MAGIC:
        raise ValueError(f"{path}: not an itnbp tensor-network file")
    (hl,) = struct.unpack("<Q", raw[8:16])
    header = json.loads(raw[16:16 + hl].decode())
    if header.get("format") != "itnbp-tensornetwork" or header.get("version") != 1:
        raise ValueError(f"{path}: unsupported format {header.get('format')!r} v{header.get('version')}")
    base = 16 + hl
    body = memoryview(raw)[base:]
    verts = [_name_from_json(v) for v in header["vertices"]]
    edges = [(verts[i], verts[j]) for i, j in header["edges"]]
    incident = {verts[i]: [edges[e][1] if edges[e][0] == verts[i] else edges[e][0] for e in inc]
                for i, inc in enumerate(header["incident"])}
    g = NamedGraph(verts, edges, incident=incident)
    dt = header["dtype"]
    tensors = {}
    for v, t in zip(verts, header["tensors"]):
        n = int(np.prod(t["shape"]))
        if dt == "bfloat16":
            bits = np.frombuffer(body, dtype="<u2", count=n, offset=t["offset"]).astype(np.uint32) << 16
            arr = bits.view(np.float32).astype(np.float64)
        elif dt == "float32":
            arr = np.frombuffer(body, dtype="<f4", count=n, offset=t["offset"]).astype(np.float32)
        else:
            arr = np.frombuffer(body, dtype="<f8", count=n, offset=t["offset"]).astype(np.float64)
        tensors[v] = arr.reshape(t["shape"], order="F")
    psi = ITensorNetwork(g, tensors)
    msgs = None
    if header.get("messages"):
        msgs = {}
        for e in header["messages"]["entries"]:
            chi = e["chi"]
            m = np.frombuffer(body, dtype="<f8", count=chi * chi, offset=e["offset"]).astype(np.float64)
            msgs[(verts[e["src"]], verts[e["dst"]])] = m.reshape(chi, chi, order="F")
    return psi, msgs
