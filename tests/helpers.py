"""Shared helpers for the tests: bridge the product's named graphs and the
oracle's integer graphs."""
import os
import re

import numpy as np

from oracle import bp_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "itensornetworks.jl_amd", "itnbp", "libitnbp.so")
HEADER = os.path.join(ROOT, "include", "itnbp.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(itnbp_\w+)\s*\(", txt, re.M)))


def oracle_graph(g):
    """oracle Graph with the same vertex numbering, edge order and axis order."""
    edges = [(g.index[a], g.index[b]) for (a, b) in g.edges]
    inc = []
    for v in g.vertices:
        inc.append([(g.index[w], g.edge_id(v, w)) for w in g.incident(v)])
    return O.Graph(g.nv, edges, incident=inc)


def oracle_psi(psi):
    return [np.asarray(psi[v], dtype=np.float64) for v in psi.graph.vertices]


def to_named(g, msgs):
    return {(g.vertices[a], g.vertices[b]): m for (a, b), m in msgs.items()}


def to_index(g, msgs):
    return {(g.index[a], g.index[b]): m for (a, b), m in msgs.items()}


def oracle_sequence(g, seq_named):
    return [(g.index[a], g.index[b]) for (a, b) in seq_named]
