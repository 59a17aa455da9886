"""Benchmark of the elimination hot path on one MI355X.

Workload (default): sequential-SLAM Schur reduction, 1000 cameras (d 6) x 200000 landmarks (d 3), track length
clip(Geometric(0.2), 2, 40) over a contiguous camera window, N(0,1) Jacobian blocks, lambda 1e-4, FP64 --
reset_schur_sparse! + reduce_schur_sparse! + copyto! of the reduced system into its full symmetric CSC.
A step is one such pass with every input resident in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
METRIC = "landmarks eliminated/s + time-to-1e-8 PCG residual, FP64 plain vs mixed/learned"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): this build has no multi-GPU path; "
              "see the sharding design in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch
    import blsload

    b = blsload.load()
    dev = b.device
    torch.cuda.set_device(0)
    t0 = time.time()
    p = b.generators.config2()
    st = p.structure
    colptr, rowval, smap = b.structure.reduced_symmetric_scatter(st, p.pattern)
    log(f"setup: {st.neblocks} landmarks, {st.nrblocks} cameras, {st.ncoup} couplings, S nnz {p.pattern.nnz}, "
        f"CSC nnz {len(smap)} ({time.time() - t0:.1f}s)")
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, b_r, b_e = T(p.E_data), T(p.C_data), T(p.b_r), T(p.b_e)
    sr, se, Rd = T(p.scaling_r), T(p.scaling_e), T(p.R_data)
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    csc = torch.empty(len(smap), dtype=torch.float64, device="cuda")
    dmap = dev.scatter_map_to_device(smap)
    lam = p.lam

    def step():
        dev.reset_schur_sparse_(S, Rd, h, sr, lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, lam, h, E, C, ws, sync=False, check=False)
        dev.copyto_(csc, S, dmap)

    # host-side check_damping once, before the timed loop (it does not change between steps)
    dev._check_damping(lam, sr, "scaling_r")
    dev._check_damping(lam, se, "scaling_e")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms = ev0.elapsed_time(ev1) / args.steps
    assert int(ws.d_failed.item()) == 0
    value = st.neblocks / (ms * 1e-3)
    log(f"timed: {args.steps} steps, {ms:.4f} ms/step (wall {wall * 1e3 / args.steps:.4f} ms/step), {value:.4g} landmarks/s")

    # per-kernel durations (HIP events recorded by the library on the launch stream)
    h.set_timing(True)
    kt = np.zeros(3)
    for _ in range(args.steps):
        dev.reset_schur_sparse_(S, Rd, h, sr, lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, lam, h, E, C, ws, check=False)
        kt += np.array(h.kernel_times())
    h.set_timing(False)
    kt /= args.steps
    k = np.diff(st.eliminated_ptr)
    nnz = p.pattern.nnz
    # algorithmic bytes (FP64): phase 2 reads E_inv (72 B), z (24 B), C (144 k B) per landmark, reads and writes S
    # once and writes g; phase 1 reads E, b_e, scaling_e and writes E_inv, z (216 B per landmark)
    p2_bytes = float(np.sum(96 + 144 * k) + 16 * nnz + 16 * st.nr)
    p1_bytes = 216.0 * st.neblocks
    whole_bytes = float(np.sum(216 + 144 * k))   # the per-landmark figure 216 + 144 k over the instance
    p2_gbs = p2_bytes / (kt[2] * 1e-3) / 1e9
    log(f"kernels: reset {kt[0]:.4f} ms, phase1 {kt[1]:.4f} ms ({p1_bytes / (kt[1] * 1e-3) / 1e9:.0f} GB/s), "
        f"phase2 {kt[2]:.4f} ms ({p2_gbs:.0f} GB/s)")
    result = {
        "metric": METRIC, "value": value, "unit": "landmarks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded PCG64 generator)",
        "config": {"workload": "reduce_schur_sparse! + copyto!, sequential SLAM, 1000 cams (d6) x 200000 landmarks (d3), "
                               "FP64", "landmarks": int(st.neblocks), "cameras": int(st.nrblocks),
                   "couplings": int(st.ncoup), "S_nnz": int(nnz), "csc_nnz": int(len(smap))},
        "kernel_ms": {"reset": kt[0], "phase1": kt[1], "phase2": kt[2]},
        "roofline": {"bound": "hbm", "kernel": "k_phase2", "achieved": p2_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": p2_gbs / HBM_PEAK_GBS, "traffic": None,
                     "step_achieved": whole_bytes / (ms * 1e-3) / 1e9},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(p, args.cpu_threads)
    print(json.dumps(result), flush=True)


def cpu_baseline(p, threads):
    """The oracle (C restatement of the reference's threaded reduction) on the host cores: reset + reduce of the
    same instance, median of repeated runs bounded to ~15 s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    st = p.structure
    threads = threads or min(16, os.cpu_count() or 1)
    # chunked like the reference's threaded path (nchunks = threads), results identical to serial
    ws = b_ws(p, threads)
    args = oracle.SchurArgs(st, ws, p.pattern)
    S = np.zeros(p.pattern.nnz)
    g = np.zeros(st.nr)
    E_inv = np.zeros(len(p.E_data))
    z = np.zeros(st.ne)
    W = np.zeros(len(p.C_data))
    times = []
    t_start = time.time()
    while len(times) < 7 and (time.time() - t_start < 15 or len(times) < 2):
        t0 = time.perf_counter()
        oracle.reset_schur_sparse(S, p.R_data, p.r_to_s, p.pattern, p.scaling_r, p.lam)
        ok, _ = oracle.reduce_schur(True, args, S, g, p.b_r, p.b_e, p.scaling_e, p.lam, p.E_data, p.C_data, E_inv, z, W,
                                    nthreads=threads)
        times.append(time.perf_counter() - t0)
        assert ok
    t = float(np.median(times))
    return {"value": st.neblocks / t, "unit": "landmarks/s", "cores": threads, "kind": "port",
            "sample": f"full workload (reset + reduce of {st.neblocks} landmarks), median of {len(times)} runs, "
                      f"C oracle -O2 with OpenMP over the reference's chunks"}


def b_ws(p, threads):
    import blsload
    b = blsload.load()
    st = p.structure
    st2 = b.structure.SchurStructure(st.r_dims, st.e_dims, (st.eliminated_ptr - 1, st.entry_retained), nchunks=threads)
    st.chunks = st2.chunks
    return b.structure.SchurReductionWorkspaceIndex(st, threads)


if __name__ == "__main__":
    main()
