"""Benchmark: steady-state FP64 CSR assembly of the global stiffness on one MI355X.

A step is one assembly of the global stiffness values (element stiffness +
deterministic scatter into the CSR pattern) for the whole mesh, with the mesh,
the pattern and the slot maps already resident in HBM.  Workload (default):
Mindlin/MITC4 plate 1024 x 1024 quads, jittered, FP64 (the plate 512 x 512,
beam frame 64^3, truss lattice 11^3 and batched workloads by --workload).

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
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PROFILES = os.path.join(ROOT, "profiles")


def profile_figure(kernel, workload, key):
    """A measured figure of `kernel` on `workload` from its committed counter
    summary profiles/<kernel>_counters__<workload>.md (a line `key: value`),
    or None when that summary does not exist."""
    path = os.path.join(PROFILES, "%s_counters__%s.md" % (kernel, workload))
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        return None
    return None


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


def backward_bytes(mesh, nnz, compliance=False):
    """Compulsory HBM bytes of one backward: dL/dK (or the 6N displacements)
    read once, geometry and parameters read once, parameter and node-coordinate
    gradients written once."""
    nn = mesh.conn.shape[1]
    n_params = {"truss": 3, "beam": 11, "plate": 4}[mesh.kind]
    src = 48 * mesh.n_nodes if compliance else 8 * nnz
    return src + 24 * mesh.n_nodes + 4 * nn * mesh.n_elements + 16 * n_params * mesh.n_elements + 24 * mesh.n_nodes


def settle(fn, stream, ms):
    """Runs fn untimed until `ms` of GPU time has passed (at most 2000 calls),
    so that the timed steps run at the clocks a sustained loop of them gets
    rather than those of a GPU that was idle during the setup."""
    import torch

    if ms <= 0:
        return 0
    done, calls = 0.0, 0
    while done < ms and calls < 2000:
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(10):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        done += ev0.elapsed_time(ev1)
        calls += 10
    return calls


def time_steps(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def sensitivity(model, mesh, nnz, val, fwd_ms, steps, warmup, stream, workload=None, settle_ms=0.0):
    """Backward from dL/dK ~ N(0, 1) on the pattern, compliance sensitivities
    with u ~ N(0, 1e-3), and forward + backward, all on the same mesh.  Each
    leg is settled like the forward (its first call also builds the adjoint
    maps on the host, during which the GPU idles)."""
    import torch

    def time_leg(fn):
        fn()
        torch.cuda.synchronize()
        settle(fn, stream, settle_ms)
        return time_steps(fn, steps, warmup, stream)

    gen = torch.Generator(device="cuda").manual_seed(0)
    g = torch.randn(nnz, dtype=torch.float64, device="cuda", generator=gen)
    u = 1e-3 * torch.randn(6 * mesh.n_nodes, dtype=torch.float64, device="cuda", generator=gen)
    xyz = mesh.kind != "beam"
    bwd_ms = time_leg(lambda: model.backward(g, xyz=xyz, stream=stream))
    comp_ms = time_leg(lambda: model.compliance_sensitivity(u, xyz=xyz, stream=stream))

    def fwd_bwd():
        model.assemble_into(val, stream)
        model.backward(g, xyz=xyz, stream=stream)

    def fwd_comp():
        model.assemble_into(val, stream)
        model.compliance_sensitivity(u, xyz=xyz, stream=stream)

    both_ms = time_leg(fwd_bwd)
    fused_ms = time_leg(lambda: model.forward_backward(val, g, xyz=xyz, stream=stream))
    fwd_comp_ms = time_leg(fwd_comp)
    bb = backward_bytes(mesh, nnz)
    cb = backward_bytes(mesh, nnz, compliance=True)
    best_ms = min(both_ms, fused_ms)
    out = {"bwd_ms": bwd_ms, "compliance_ms": comp_ms, "fwd_then_bwd_ms": both_ms,
            "fwd_bwd_two_streams_ms": fused_ms, "fwd_plus_bwd_ms": best_ms,
            "fwd_plus_bwd_over_fwd": best_ms / fwd_ms,
            # forward then the compliance sensitivities -u_e^T (dK_e/dp) u_e
            # (strain-energy form, no dL/dK read)
            "fwd_then_compliance_ms": fwd_comp_ms, "fwd_plus_compliance_over_fwd": fwd_comp_ms / fwd_ms,
            "sensitivities_per_s": mesh.n_elements / (best_ms * 1e-3),
            "bwd_roofline": {"bound": "hbm", "achieved": bb / (bwd_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": bb / (bwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "algorithmic_bytes_per_launch": bb},
            "compliance_roofline": {"bound": "hbm", "achieved": cb / (comp_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                                    "unit": "GB/s", "frac": cb / (comp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                    "algorithmic_bytes_per_launch": cb}}
    # FP64 arithmetic of the plate backward and compliance launches (committed
    # counter summary) against the FP64 peak, at the measured step times
    for key, ms in (("bwd", bwd_ms), ("compliance", comp_ms)):
        fl = profile_figure("plate_adjoint", workload, key + "_fp64_flops_per_launch") if workload else None
        if fl is not None:
            tf = fl / (ms * 1e-3) / 1e12
            out[key + "_fp64"] = {"flops_per_launch": fl, "achieved_tflops": tf, "peak_tflops": FP64_PEAK_TFLOPS,
                                  "frac": tf / FP64_PEAK_TFLOPS}
    return out


def separation(model, mesh, nnz, steps, warmup):
    """Sparse separation (K_aa, K_ab, K_ba, K_bb, k_aa/k_bb indexes, r_a, u_b)
    of the assembled K with the x = min face restrained (plates: Z, ThX, ThY;
    trusses: X, Y, Z; beams: all six dofs) and a load on every other node's Z.
    Kernel time from HIP events around the separation kernels; the host read
    of the result sizes is outside it."""
    x = mesh.xyz[:, 0]
    face = np.nonzero(np.isclose(x, x.min()))[0]
    dofs = {"plate": (2, 3, 4), "truss": (0, 1, 2), "beam": tuple(range(6))}[mesh.kind]
    numbers = np.repeat(mesh.node_numbers[face], len(dofs))
    model.add_displacements(numbers, np.tile(dofs, len(face)), np.zeros(len(numbers)))
    rest = np.setdiff1d(np.arange(mesh.n_nodes), face)
    model.add_concentrated_loads(mesh.node_numbers[rest], np.full(len(rest), 2), np.full(len(rest), -1.0))
    n_aa, n_bb, nnz_sep = model.separate(assemble=True)  # uploads the boundary conditions
    for _ in range(warmup):
        model.separate(assemble=False)
    t0 = time.perf_counter()
    kern = 0.0
    for _ in range(steps):
        model.separate(assemble=False)
        kern += model.last_separate_ms()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    kern /= steps
    n = 6 * mesh.n_nodes
    kept = sum(nnz_sep.values())
    # compulsory bytes: the CSR read once, the four blocks, their row pointers,
    # the index vectors and r_a / u_b written once, the restraint flags read once
    alg = 12 * nnz + 8 * (n + 1) + n + 12 * kept + 8 * (2 * (n_aa + n_bb) + 4) + 16 * (n_aa + n_bb)
    return {"separate_ms": kern, "separate_wall_ms": wall_ms, "n_aa": n_aa, "n_bb": n_bb, "nnz": nnz_sep,
            "roofline": {"bound": "hbm", "achieved": alg / (kern * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": alg / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_launch": alg}}


def analysis(model, mesh, sep, steps, warmup, pcg_iters):
    """Global analysis on the separation above: block-Jacobi PCG capped at
    `pcg_iters` updates (cost per update; the rate is what matters, not
    convergence), then element forces from the resulting displacements.
    PCG time is the HIP-event time of the whole solve, including the host
    read of r.r after every update; forces are timed with events on the
    launch stream."""
    import torch

    n_aa, nnz_aa = sep["n_aa"], sep["nnz"]["aa"]
    model.solve_pcg("block_jacobi", max_iter=pcg_iters)
    ms, iters = [], 0
    for _ in range(3):
        iters = model.solve_pcg("block_jacobi", max_iter=pcg_iters)[2]
        ms.append(model.last_solve_ms())
    per_it = min(ms) / max(iters, 1)
    # compulsory bytes of one update of a fused PCG: K_aa (values, columns, row
    # pointers) once; p read and K p written by the product; x, p, r, K p read
    # and x, r, z written by the update; z, p read and p written; the node
    # blocks of the preconditioner (36 doubles per 6 dofs) read once
    spmv = 12 * nnz_aa + 8 * (n_aa + 1)
    it_bytes = spmv + 8 * 12 * n_aa + 48 * n_aa
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        model.element_forces(stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        model.element_forces(stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    f_ms = e0.elapsed_time(e1) / steps
    nn = mesh.conn.shape[1]
    n_params = {"truss": 3, "beam": 11, "plate": 4}[mesh.kind]
    n_out = {"truss": 1, "beam": 10, "plate": 8}[mesh.kind]
    # displacements once, connectivity, parameters, frames and forces once
    f_bytes = 48 * mesh.n_nodes + mesh.n_elements * (4 * nn + 8 * n_params + 72 + 8 * n_out)

    def roof(b, t_ms):
        a = b / (t_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                "algorithmic_bytes_per_launch": b}

    return {"pcg": {"preconditioner": "block_jacobi", "updates": iters, "ms_per_update": per_it,
                    "updates_per_s": 1e3 / per_it, "roofline": roof(it_bytes, per_it)},
            "element_forces": {"ms": f_ms, "elements_per_s": mesh.n_elements / (f_ms * 1e-3),
                               "roofline": roof(f_bytes, f_ms)}}


def run_batched(args):
    """Batched workload: 256 independent meshes (type uniform over truss
    lattice / beam frame / plate, element counts log-uniform in [1e3, 2e5]),
    assembled by one graph launch vs a loop of per-mesh launches."""
    import ctypes

    import torch

    from finite_element_method_amd import FEM, FEMBatch, lib, meshgen

    meshes = meshgen.batched_meshes(256, seed=0)
    models = [FEM.from_mesh(m) for m in meshes]
    t0 = time.perf_counter()
    vals, nnz_total, alg = [], 0, 0
    for mesh, model in zip(meshes, models):
        model.build_pattern()
        nnz = model.csr_info()[1]
        nnz_total += nnz
        alg += algorithmic_bytes(mesh, nnz)
        vals.append(torch.empty(nnz, dtype=torch.float64, device="cuda:0"))
    batch = FEMBatch(models)
    config = batch.autotune()
    setup_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream()
    handles = [(m._h, ctypes.c_void_p(v.data_ptr())) for m, v in zip(models, vals)]
    sh = ctypes.c_void_p(stream.cuda_stream)
    L = lib()

    def loop():
        for h, v in handles:
            L.fem_assemble_into(h, v, sh)

    settle(loop, stream, args.settle_ms)
    loop_ms = time_steps(loop, args.steps, args.warmup, stream)
    settled = settle(lambda: batch.assemble(stream), stream, args.settle_ms)
    t_wall = time.perf_counter()
    batch_ms = time_steps(lambda: batch.assemble(stream), args.steps, args.warmup, stream)
    wall = time.perf_counter() - t_wall
    n_el = sum(m.n_elements for m in meshes)
    achieved = alg / (batch_ms * 1e-3) / 1e9
    kinds = {k: sum(1 for m in meshes if m.kind == k) for k in ("truss", "beam", "plate")}
    out = {
        "metric": METRIC, "value": n_el / (batch_ms * 1e-3), "unit": "elements/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": batch_ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "batched_256", "meshes": len(meshes), "kinds": kinds, "elements": n_el,
                   "nnz": nnz_total, "seed": 0, "launch": "one hipGraph launch", "tuned": config},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes_per_launch": alg},
        "per_mesh_loop_ms": loop_ms, "speedup_vs_loop": loop_ms / batch_ms,
        "setup_s": setup_s, "wall_s": wall, "settle": {"ms": args.settle_ms, "calls": settled},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline("plate")
    print(json.dumps(out), flush=True)


def cpu_baseline(kind):
    """The oracle (the crate's algorithm restated: dense element k, (T^T k) T,
    insertion-order sums into the CSR) on a bounded sample of the same workload
    family: the steady-state value recompute timed single-threaded and on all
    OpenMP threads of the host share (results checked bitwise against the
    sequential export)."""
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
    om = meshgen.load_oracle(sample, OracleModel)
    om.csr(structural=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    one, bad1 = om.time_values(1, reps=5, warmup=1)
    par, badp = om.time_values(threads, reps=5, warmup=1)
    if bad1 or badp:
        raise RuntimeError("CPU baseline values differ from the sequential export")
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": sample.n_elements / par, "unit": "elements/s", "cores": threads, "kind": "port",
            "single_thread_value": sample.n_elements / one, "cpu_model": model, "nproc": os.cpu_count(),
            "sample": desc + ", oracle value recompute (element k, T^T k T, insertion-order CSR sums), "
                      "median of 5 after 1 warm-up on %d OpenMP threads (%.3f s) and on one thread "
                      "(%.3f s)" % (threads, par, one)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle-ms", type=float, default=100.0,
                    help="untimed GPU time spent on the step before the warm-up steps (clocks settle)")
    ap.add_argument("--workload", default="plate_1024", choices=sorted(WORKLOADS) + ["batched_256"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-grad", action="store_true", help="skip the backward / sensitivity timings")
    ap.add_argument("--no-separate", action="store_true", help="skip the sparse separation timing")
    ap.add_argument("--pcg-iters", type=int, default=50,
                    help="PCG updates per timed solve after the separation (0 = skip the analysis leg)")
    ap.add_argument("--separate-lanes", type=int, default=0, help="separation: lanes per CSR row (0 = automatic)")
    ap.add_argument("--plate-kernel", type=int, default=-1, help="fused plate kernel: 0 = row staging, 1 = block lanes, 2 = corner lanes")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one elementwise copy of this many MiB first (a known-traffic kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    ap.add_argument("--patch-nodes", type=int, default=0, help="fused plate: owned nodes per patch (0 = default)")
    ap.add_argument("--wg-per-cu", type=int, default=-1, help="fused plate: resident workgroups per CU (0 = one per patch)")
    ap.add_argument("--assembly-events", type=int, default=0,
                    help="1: the library records its own events around every assembly launch (fem_last_assembly_ms); "
                         "the timed loop is bracketed by the bench's events either way")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(multi-GPU sharding is a design note, not implemented)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from finite_element_method_amd import FEM, meshgen

    if args.workload == "batched_256":
        run_batched(args)
        return
    if args.copy_probe_mib:
        n = args.copy_probe_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    gen, kw = WORKLOADS[args.workload]
    mesh = getattr(meshgen, gen)(**kw)
    model = FEM.from_mesh(mesh)
    if args.patch_nodes:
        model.set_option("plate_patch_nodes", args.patch_nodes)
    if args.wg_per_cu >= 0:
        model.set_option("plate_wg_per_cu", args.wg_per_cu)
    if args.separate_lanes:
        model.set_option("separate_lanes", args.separate_lanes)
    if args.plate_kernel >= 0:
        model.set_option("plate_kernel", args.plate_kernel)
    model.set_option("time_assembly", args.assembly_events)
    t0 = time.perf_counter()
    model.build_pattern()
    setup_s = time.perf_counter() - t0
    n_rows, nnz = model.csr_info()
    dev = torch.device("cuda:0")
    val = torch.empty(nnz, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    settled = settle(lambda: model.assemble_into(val, stream), stream, args.settle_ms)
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
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profile_figure(model.assembly_path(), args.workload, "traffic_bytes_per_launch"),
                     "algorithmic_bytes_per_launch": alg},
        "setup_s": setup_s,
        "wall_s": wall,
        "settle": {"ms": args.settle_ms, "calls": settled},
    }
    flops = profile_figure(model.assembly_path(), args.workload, "fp64_flops_per_launch")
    if flops is not None:
        # FP64 arithmetic of the same launch (SQ_INSTS_VALU_FLOPS_FP64 and the
        # F64 MFMA count of the committed counter summary) against the FP64 peak
        mfma = profile_figure(model.assembly_path(), args.workload, "fp64_mfma_flops_per_launch") or 0.0
        out["fp64"] = {"flops_per_launch": flops, "achieved_tflops": flops / (ms * 1e-3) / 1e12,
                       "peak_tflops": FP64_PEAK_TFLOPS, "frac": flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                       "mfma_flops_per_launch": mfma, "mfma_frac": mfma / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    if not args.no_grad:
        out["sensitivity"] = sensitivity(model, mesh, nnz, val, ms, args.steps, args.warmup, stream, args.workload,
                                         args.settle_ms)
    if not args.no_separate:
        out["separation"] = separation(model, mesh, nnz, args.steps, args.warmup)
        if args.pcg_iters > 0:
            out["analysis"] = analysis(model, mesh, out["separation"], args.steps, args.warmup, args.pcg_iters)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh.kind)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
