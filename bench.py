"""Benchmark: steady-state FP64 CSR assembly of the global stiffness on one MI355X.

A step is one assembly of the global stiffness values (element stiffness +
deterministic scatter into the CSR pattern) for the whole mesh, with the mesh,
the pattern and the slot maps already resident in HBM.  Workload (default):
Mindlin/MITC4 plate 512 x 512 quads, jittered, FP64.

Prints ONE JSON line.  --gpus above 1 exits with a message (single-GPU build).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 elements assembled/s + fwd+bwd sensitivities/s; % of HBM roofline vs CPU"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "plate_512": ("plate_grid", dict(nx=512, seed=0)),
    "plate_1024": ("plate_grid", dict(nx=1024, seed=0)),
    "beam_frame_64": ("beam_frame", dict(nx=64, seed=0)),
    "truss_lattice_11": ("truss_lattice", dict(n=11, seed=0)),
}


def algorithmic_bytes(mesh, nnz):
    """Compulsory HBM bytes of one steady-state assembly: values written once,
    node coordinates, connectivity and per-element parameters read once."""
    nn = mesh.conn.shape[1]
    n_params = {"truss": 3, "beam": 11, "plate": 4}[mesh.kind]
    return 8 * nnz + 24 * mesh.n_nodes + 4 * nn * mesh.n_elements + 8 * n_params * mesh.n_elements


def cpu_baseline(kind):
    """The oracle (single-threaded restatement of the crate's algorithm: dense
    element k, (T^T k) T, sequential accumulation, CSR export) on a bounded
    sample of the same workload family."""
    from finite_element_method_amd import meshgen
    from oracle.oracle import OracleModel

    if kind == "plate":
        sample = meshgen.plate_grid(320, seed=0)
        desc = "plate_grid(320): 102400 jittered MITC4 quads"
    elif kind == "beam":
        sample = meshgen.beam_frame(24, seed=0)
        desc = "beam_frame(24): 39744 beams"
    else:
        sample = meshgen.truss_lattice(11, seed=0)
        desc = "truss_lattice(11): full workload"
    t0 = time.perf_counter()
    om = meshgen.load_oracle(sample, OracleModel)
    om.csr(structural=True)
    dt = time.perf_counter() - t0
    return {"value": sample.n_elements / dt, "unit": "elements/s", "cores": 1, "kind": "port",
            "sample": desc + ", oracle add_* (validation, element k, T^T k T, scatter) + CSR export, %.1f s" % dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="plate_512", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--patch-nodes", type=int, default=0, help="fused plate: owned nodes per patch (0 = default)")
    ap.add_argument("--wg-per-cu", type=int, default=-1, help="fused plate: resident workgroups per CU (0 = one per patch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is a design note, not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from finite_element_method_amd import FEM, meshgen

    gen, kw = WORKLOADS[args.workload]
    mesh = getattr(meshgen, gen)(**kw)
    model = FEM.from_mesh(mesh)
    if args.patch_nodes:
        model.set_option("plate_patch_nodes", args.patch_nodes)
    if args.wg_per_cu >= 0:
        model.set_option("plate_wg_per_cu", args.wg_per_cu)
    t0 = time.perf_counter()
    model.build_pattern()
    setup_s = time.perf_counter() - t0
    n_rows, nnz = model.csr_info()
    dev = torch.device("cuda:0")
    val = torch.empty(nnz, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        model.assemble_into(val, stream)
    torch.cuda.synchronize()

    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        model.assemble_into(val, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = ev0.elapsed_time(ev1) / args.steps

    elems_per_s = mesh.n_elements / (ms * 1e-3)
    alg = algorithmic_bytes(mesh, nnz)
    achieved = alg / (ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": elems_per_s,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "elements": mesh.n_elements, "nodes": mesh.n_nodes,
                   "dof": int(n_rows), "nnz": int(nnz), "seed": 0, "path": model.assembly_path()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": alg},
        "setup_s": setup_s,
        "wall_s": wall,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh.kind)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
