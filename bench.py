"""Benchmark of the elimination hot path on one MI355X.

Workload (default): sequential-SLAM Schur reduction, 1000 cameras (d 6) x 200000 landmarks (d 3), track length
clip(Geometric(0.2), 2, 40) over a contiguous camera window, N(0,1) Jacobian blocks, lambda 1e-4, FP64 --
reset_schur_sparse! + reduce_schur_sparse! + copyto! of the reduced system into its full symmetric CSC.
A step is one such pass with every input resident in HBM. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--calibration-copy]

--calibration-copy first runs one device copy of 256 MiB (a kernel of known traffic) so that a hardware-counter
run of this command can convert FETCH_SIZE / WRITE_SIZE into bytes. roofline.traffic is read from the committed
counter summary profiles/<phase-2 kernel>_slam_traffic.md when present.
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
FP64_MFMA_PEAK_TF = 78.6


def measured_traffic(kernel):
    """HBM bytes per launch of the phase-2 kernel from its counter summary (line `traffic_bytes_per_launch: N`)."""
    try:
        with open(os.path.join(ROOT, "profiles", f"{kernel}_slam_traffic.md")) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibration-copy", action="store_true")
    ap.add_argument("--qr-scale", type=float, default=1.0,
                    help="landmark scale of the reduce_qr! leg (500 cameras d 6/9, 500000 landmarks x scale); 0 skips")
    ap.add_argument("--config3-scale", type=float, default=1.0,
                    help="landmark scale of the dense leg (256 cameras, 1000000 landmarks x scale); 0 skips it")
    ap.add_argument("--config3-valu-scale", type=float, default=1.0,
                    help="fraction of the dense leg's landmarks at which the VALU kernel, the CPU oracle and the "
                         "parity check run (1: the same full-size instance as the MFMA path)")
    ap.add_argument("--config5-scale", type=float, default=1.0,
                    help="landmark scale of the batched 64-problem leg (cameras at full size); 0 skips it")
    ap.add_argument("--tuning", default="",
                    help="PART,LANES: use this phase-2 work-list shape instead of autotuning (profiling runs)")
    ap.add_argument("--slam-only", action="store_true",
                    help="time only the sequential-SLAM step (no other legs, no CPU baseline): for profiler runs")
    ap.add_argument("--no-dense-solve", action="store_true",
                    help="skip the dense reduced solve leg (its Bunch-Kaufman fallback launches two kernels per "
                         "column): for hardware-counter runs of the other legs")
    ap.add_argument("--config3-only", action="store_true",
                    help="run only the dense leg (no SLAM line, no CPU baseline): for profiler runs of its kernels")
    args = ap.parse_args()
    if args.slam_only:
        args.no_cpu, args.qr_scale, args.config5_scale, args.config3_scale = True, 0.0, 0.0, 0.0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the library has no multi-GPU path "
              "(the reduction shards by independent problems or by landmark ranges, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import blsload

    b = blsload.load()
    dev = b.device
    torch.cuda.set_device(0)
    if args.config3_only:
        print(json.dumps({"dense_reduction": config3_leg(torch, dev, b, args.config3_scale or 1.0,
                                                         valu_scale=args.config3_valu_scale, cpu=False)}),
              flush=True)
        return
    if args.calibration_copy:
        src = torch.ones(256 << 17, dtype=torch.float64, device="cuda")   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
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
    # the autotuner picks the phase-2 work-list shape on this instance (landmarks per part, rows per run)
    t0 = time.time()
    cands = [(pt, ln) for pt in (384, 768, 1152, 1536, 2304) for ln in (32, 48, 64, 96)]
    if args.tuning:
        cands = [tuple(int(v) for v in args.tuning.split(","))]
        h.set_tuning(*cands[0])
        tuned = cands[0] + (float("nan"),)
    else:
        tuned = h.autotune(cands, S, g, b_r, b_e, se, lam, E, C, ws)
    info = h.phase2_info()
    p2_kernel = "k_tiles" if info[0] == 2 else "k_phase2"
    log(f"autotune: part {tuned[0]}, lanes {tuned[1]} ({tuned[2]:.4f} ms reduce; default 384 / 96) "
        f"over {len(cands)} shapes and the tiles ({time.time() - t0:.1f}s); phase 2 runs {p2_kernel} "
        f"(tile plan: {info[1]} tiles, {info[2]} landmark entries, {info[3]} matrix-core products)")
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
    eager_ms = ev0.elapsed_time(ev1) / args.steps
    assert int(ws.d_failed.item()) == 0
    log(f"eager: {args.steps} steps, {eager_ms:.4f} ms/step (wall {wall * 1e3 / args.steps:.4f} ms/step), "
        f"{st.neblocks / (eager_ms * 1e-3):.4g} landmarks/s")
    # the same step captured once in a HIP graph and replayed (what a solver loop over one structure runs): the
    # kernels and their work are identical, the launch overhead is gone
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        graph.replay()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    graph_ms = ev0.elapsed_time(ev1) / args.steps
    assert int(ws.d_failed.item()) == 0
    log(f"graph replay: {args.steps} steps, {graph_ms:.4f} ms/step (wall {wall * 1e3 / args.steps:.4f} ms/step), "
        f"{st.neblocks / (graph_ms * 1e-3):.4g} landmarks/s")
    del graph
    ms = eager_ms   # the metric: the library's calls launched as a caller makes them
    value = st.neblocks / (ms * 1e-3)

    # per-kernel durations (HIP events recorded by the library on the launch stream)
    h.set_timing(True)
    kt = np.zeros(4)
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
        f"phase2 {kt[2]:.4f} ms ({p2_gbs:.0f} GB/s), merge {kt[3]:.4f} ms")
    # backsubstitute! (x_e = z - E_inv sum_c C_c x_r) from the reduction just run: C_data, E_inv and z read once, x_r
    # gathered, x_e written
    xr = torch.randn(st.nr, dtype=torch.float64, device="cuda")
    xe = torch.empty(st.ne, dtype=torch.float64, device="cuda")
    bms = time_ms(torch, lambda: dev.backsubstitute_(xe, xr, h, C, ws), 10, batch=10)
    bbytes = 8.0 * (p.C_data.size + st.eliminated_storage_length() + 2 * st.ne + st.nr)
    log(f"backsubstitute!: {bms:.4f} ms, {bbytes / (bms * 1e-3) / 1e9:.0f} GB/s of {bbytes / 1e9:.3f} GB")
    backsub = {"ms": bms, "achieved_GBs": bbytes / (bms * 1e-3) / 1e9, "bytes": bbytes}
    mixed = assembly = pcie = None
    if not args.slam_only:
        mixed = mixed_leg(torch, dev, p, h, S, csc, dmap, args.steps, args.warmup)
        assembly = assembly_leg(torch, dev, st)
        pcie = pcie_inclusive(torch, dev, p, h, ws, S, g, csc, dmap, max(3, args.steps // 4))
        log(f"PCIe-inclusive (host E, C, b, R in; S, g, CSC out): {pcie['ms_per_step']:.3f} ms/step, "
            f"{pcie['value']:.4g} landmarks/s")
    result = {
        "metric": METRIC, "value": value, "unit": "landmarks/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded PCG64 generator)",
        "graph_ms_per_step": graph_ms, "step": "reset + reduce + copyto!, device-resident inputs",
        "tuning": {"part": tuned[0], "lanes": tuned[1], "candidates": len(cands), "phase2_kernel": p2_kernel,
                   "tiles": info[1], "tile_entries": info[2], "tile_mfma": info[3]},
        "config": {"workload": "reduce_schur_sparse! + copyto!, sequential SLAM, 1000 cams (d6) x 200000 landmarks (d3), "
                               "FP64", "landmarks": int(st.neblocks), "cameras": int(st.nrblocks),
                   "couplings": int(st.ncoup), "S_nnz": int(nnz), "csc_nnz": int(len(smap))},
        "kernel_ms": {"reset": kt[0], "phase1": kt[1], "phase2": kt[2], "phase2_merge": kt[3]},
        "roofline": {"bound": "hbm", "kernel": p2_kernel, "achieved": p2_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": p2_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(p2_kernel),
                     "step_achieved": whole_bytes / (ms * 1e-3) / 1e9,
                     # the tiles' issued FP64 matrix-core rate (16 x 16 x 4 products, K padded from d_e)
                     "mfma_issued_TFs": (info[3] * 2048.0 / (kt[2] * 1e-3) / 1e12) if info[0] == 2 else None,
                     "mfma_peak_TFs": FP64_MFMA_PEAK_TF},
        "pcie_inclusive": pcie,
        "mixed": mixed,
        "assembly": assembly,
        "backsubstitute": backsub,
    }
    # the step's outputs, kept for the full-size parity check against the oracle
    S_dev, g_dev = S.cpu().numpy(), g.cpu().numpy()
    del E, C, b_r, b_e, Rd, S, g, csc, dmap, h, ws
    torch.cuda.empty_cache()
    if not args.slam_only and not args.no_dense_solve:
        result["dense_cholesky"] = dense_leg(torch, dev, int(st.nr))
    torch.cuda.empty_cache()
    if args.qr_scale > 0:
        result["reduce_qr"] = qr_leg(torch, dev, b, args.qr_scale)
        torch.cuda.empty_cache()
    if args.config3_scale > 0:
        result["dense_reduction"] = config3_leg(torch, dev, b, args.config3_scale, valu_scale=args.config3_valu_scale,
                                                cpu=not args.no_cpu)
        torch.cuda.empty_cache()
    if args.config5_scale > 0:
        result["batched_pcg"] = batched_pcg_leg(torch, dev, b, args.config5_scale, check_parity=not args.no_cpu)
    if not args.no_cpu:
        result["cpu_baseline"], S_ref, g_ref = cpu_baseline(p, args.cpu_threads)
        result["parity"] = parity_record(st, p.pattern, S_dev, g_dev, S_ref, g_ref)
        log(f"parity at full size against the oracle: S max block rel {result['parity']['max_block_rel']:.3g}, "
            f"g max block rel {result['parity']['g_max_block_rel']:.3g}")
    print(json.dumps(result), flush=True)
    checks = [result.get("parity"), (result.get("batched_pcg") or {}).get("parity")]
    if any(c is not None and not c["ok"] for c in checks):
        sys.exit("parity check failed: the device's S or g is more than 1e-12 (relative Frobenius per block) from "
                 "the oracle's")


PARITY_TOL = 1e-12


def sparse_block_rel(st, pattern, A, B):
    """Largest ||A_b - B_b||_F / ||B_b||_F over the upper blocks b of a sparse S (only the upper triangle of a
    diagonal block is compared; absolute where B_b = 0)."""
    nr, nrb = st.nr, st.nrblocks
    colptr = np.asarray(pattern.S_colptr) - 1
    rows = np.asarray(pattern.S_rowval) - 1
    cols = np.repeat(np.arange(nr), np.diff(colptr))
    bo = np.asarray(st.boffsets_r) - 1
    brow = np.searchsorted(bo, rows, side="right") - 1
    bcol = np.searchsorted(bo, cols, side="right") - 1
    keep = ~((brow == bcol) & (rows > cols))
    key = (brow * nrb + bcol)[keep]
    _, blk = np.unique(key, return_inverse=True)
    d2 = np.bincount(blk, weights=(A[keep] - B[keep]) ** 2)
    n2 = np.bincount(blk, weights=B[keep] ** 2)
    rel = np.where(n2 > 0, np.sqrt(d2) / np.sqrt(np.where(n2 > 0, n2, 1.0)), np.sqrt(d2))
    return float(rel.max())


def vector_block_rel(dims, a, b):
    """Largest relative 2-norm error over the blocks (of sizes `dims`) of a vector."""
    blk = np.repeat(np.arange(len(dims)), dims)
    d2 = np.bincount(blk, weights=(a - b) ** 2, minlength=len(dims))
    n2 = np.bincount(blk, weights=b ** 2, minlength=len(dims))
    rel = np.where(n2 > 0, np.sqrt(d2) / np.sqrt(np.where(n2 > 0, n2, 1.0)), np.sqrt(d2))
    return float(rel.max())


def parity_record(st, pattern, S_dev, g_dev, S_ref, g_ref):
    s_rel = sparse_block_rel(st, pattern, S_dev, S_ref)
    g_rel = vector_block_rel(np.asarray(st.r_dims), g_dev, g_ref)
    return {"max_block_rel": s_rel, "g_max_block_rel": g_rel,
            "g_rel": float(np.linalg.norm(g_dev - g_ref) / np.linalg.norm(g_ref)), "tol": PARITY_TOL,
            "ok": bool(s_rel <= PARITY_TOL and g_rel <= PARITY_TOL),
            "against": "the oracle's Float64 reset + reduce of the same instance"}


FP32_MFMA_PEAK_TF = 157.3


def config3_leg(torch, dev, b, scale, valu_scale=1.0, cpu=True, reps=3):
    """The dense configuration: 256 cameras (d 6), 1000000 x scale landmarks (d 3), k ~ U{2..64} random cameras,
    Jacobian blocks scaled per camera by 10^U(-2, 2), lambda 1e-4, dense reduce_schur_dense!.
    - The SYRK on the matrix cores (phase 1 in Float64, S -= Y'Y with Y_e = chol(E_inv_e)' C_e): Float64 C and FP64
      MFMA, and Float32 couplings with FP32 MFMA and Float64 cross-slice accumulation; their dense product's
      MFMA rate against the peaks, and the useful rate (108 k^2 + 162 k + 68 flop per landmark).
    - The mixed S against the Float64 one, and the dense solve through the mixed S's Cholesky factor with one
      Float64 refinement step (its residual from the Float64 blocks), judged against the Float64 system.
    - The owner-computes VALU kernel (the default dense reduction) at valu_scale of the landmarks (the full instance
      by default; the MFMA path is timed again there when it is smaller), the CPU oracle on the same instance and the
      VALU result's parity against it."""
    t0 = time.time()
    p = b.generators.config3(scale=scale)
    st = p.structure
    log(f"dense leg setup: {st.neblocks} landmarks, {st.nrblocks} cameras, S {st.nr} x {st.nr} ({time.time() - t0:.1f}s)")
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    k = np.diff(st.eliminated_ptr).astype(np.float64)
    useful = float(np.sum(108 * k * k + 162 * k + 68))
    nr = st.nr
    h = dev.SchurHandle(st, p.ws_index, dense=True)
    ws = dev.SchurReductionWorkspace(h)
    Rv, sr, br, be, se, E = T(p.R.T.reshape(-1)), T(p.scaling_r), T(p.b_r), T(p.b_e), T(p.scaling_e), T(p.E_data)
    out = {"landmarks": int(st.neblocks), "cameras": int(st.nrblocks), "useful_flop": useful}
    Ss, gs = {}, {}
    ntiles = (nr + 127) // 128
    dense_flop = 2.0 * st.ne * (ntiles * (ntiles + 1) // 2) * 128 * 128
    for name, dt, peak in (("mfma_f64", torch.float64, FP64_MFMA_PEAK_TF), ("mfma_mixed", torch.float32, FP32_MFMA_PEAK_TF)):
        C = T(p.C_data, dt)
        S = torch.empty(nr * nr, dtype=torch.float64, device="cuda")
        g = torch.empty(nr, dtype=torch.float64, device="cuda")
        ts = []
        for rep_ in range(reps + 1):
            dev.reset_schur_dense_(S, Rv, h, sr, p.lam, check=False)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            assert dev.reduce_schur_dense_mfma_(S, g, br, be, se, p.lam, h, E, C, ws, check=False)
            ts.append(time.perf_counter() - t1)
        h.set_timing(True)
        dev.reset_schur_dense_(S, Rv, h, sr, p.lam, check=False)
        dev.reduce_schur_dense_mfma_(S, g, br, be, se, p.lam, h, E, C, ws, check=False)
        kt = h.kernel_times()
        h.set_timing(False)
        ms = float(np.median(ts[1:])) * 1e3
        mfma_tf = dense_flop / (kt[2] * 1e-3) / 1e12
        out[name] = {"ms": ms, "landmarks_per_s": st.neblocks / (ms * 1e-3), "useful_TFLOPs": useful / (ms * 1e-3) / 1e12,
                     "kernel_ms": {"ybuild": kt[1], "syrk": kt[2], "slice_sum": kt[3]},
                     "roofline": {"bound": "mfma", "kernel": "k_syrk", "achieved": mfma_tf, "peak": peak,
                                  "unit": "TFLOP/s", "frac": mfma_tf / peak,
                                  "dense_flop": dense_flop, "useful_fraction_of_dense": useful / dense_flop}}
        log(f"dense leg, {name}: {ms:.2f} ms ({st.neblocks / (ms * 1e-3):.4g} landmarks/s, {useful / (ms * 1e-3) / 1e12:.2f} "
            f"useful TFLOP/s); SYRK {kt[2]:.2f} ms = {mfma_tf:.1f} TFLOP/s of the dense product ({100 * mfma_tf / peak:.0f}% "
            f"of {peak}), Y build {kt[1]:.2f} ms")
        Ss[name], gs[name] = S, g
        del C
    S64, g64, Sm, gm = Ss["mfma_f64"], gs["mfma_f64"], Ss["mfma_mixed"], gs["mfma_mixed"]
    A64 = S64.view(nr, nr).T.triu()
    Am = Sm.view(nr, nr).T.triu()
    out["mfma_mixed"]["rel_diff_S_vs_f64"] = float((Am - A64).norm() / A64.norm())
    out["mfma_mixed"]["rel_diff_g_vs_f64"] = float((gm - g64).norm() / g64.norm())
    # the dense solve through the mixed S with one Float64 refinement step, against the Float64 system
    R = b.refine
    SR = torch.empty_like(S64)
    dev.reset_schur_dense_(SR, Rv, h, sr, p.lam, check=False)
    rd = R.RefinementData(h, E, T(p.C_data), br, be, se, SR, p.lam)
    chol = dev.DenseCholesky(nr)
    x = torch.empty(nr, dtype=torch.float64, device="cuda")
    r = torch.empty_like(x)

    def relres_f64(xv):
        dev.dense_residual_(r, S64, g64, xv)
        return float(r.norm() / g64.norm())

    torch.cuda.synchronize()
    t1 = time.perf_counter()
    assert R.factor(rd, check=False)
    ok, _ = chol.factor(Sm)
    assert ok
    R.refined_dense_solve(chol, gm, x, rd, steps=1)
    torch.cuda.synchronize()
    solve_ms = (time.perf_counter() - t1) * 1e3
    refined = relres_f64(x)
    R.refined_dense_solve(chol, gm, x, rd, steps=0)
    unrefined = relres_f64(x)
    out["mfma_mixed"]["refined_solve"] = {"ms": solve_ms, "relres_vs_f64_system_unrefined": unrefined,
                                          "relres_vs_f64_system_one_step": refined,
                                          "time_to_refined_solution_ms": out["mfma_mixed"]["ms"] + solve_ms}
    log(f"dense leg, mixed: ||S_mixed - S_f64|| / ||S_f64|| = {out['mfma_mixed']['rel_diff_S_vs_f64']:.2e}; the solve "
        f"through its Cholesky factor: residual against the Float64 system {unrefined:.2e}, {refined:.2e} after one "
        f"Float64 refinement step ({solve_ms:.2f} ms incl. the Float64 factorisation)")
    del Ss, gs, S64, Sm, A64, Am, SR, rd, E, h, ws
    torch.cuda.empty_cache()
    # the VALU owner-computes kernel, the MFMA path and the CPU oracle at valu_scale of the landmarks
    q = p if valu_scale == 1.0 else b.generators.config3(scale=scale * valu_scale)
    sq = q.structure
    t0 = time.time()
    hq = dev.SchurHandle(sq, q.ws_index, dense=True)
    plan_s = time.time() - t0
    wq = dev.SchurReductionWorkspace(hq)
    Eq, Cq, Rq = T(q.E_data), T(q.C_data), T(q.R.T.reshape(-1))
    brq, beq, srq, seq = T(q.b_r), T(q.b_e), T(q.scaling_r), T(q.scaling_e)
    Sq = torch.empty(sq.nr * sq.nr, dtype=torch.float64, device="cuda")
    gq = torch.empty(sq.nr, dtype=torch.float64, device="cuda")
    kq = np.diff(sq.eliminated_ptr).astype(np.float64)
    useful_q = float(np.sum(108 * kq * kq + 162 * kq + 68))
    small = {"landmarks": int(sq.neblocks), "valu_work_list_s": plan_s}
    # the work-list shape on this instance (columns of 256 block rows: wider items stage each landmark fewer times)
    t0 = time.time()
    dev.reset_schur_dense_(Sq, Rq, hq, srq, q.lam, check=False)
    tuned = hq.autotune([(384, 96), (1536, 96), (1536, 192), (1536, 256)], Sq, gq, brq, beq, seq, q.lam, Eq, Cq, wq)
    small["tuning"] = {"part": tuned[0], "lanes": tuned[1]}
    log(f"dense leg autotune: part {tuned[0]}, lanes {tuned[1]} ({tuned[2]:.3f} ms reduce) ({time.time() - t0:.1f}s)")
    legs = [("valu_f64", lambda: dev.reduce_schur_dense_(Sq, gq, brq, beq, seq, q.lam, hq, Eq, Cq, wq, check=False))]
    if valu_scale != 1.0:
        legs.append(("mfma_f64", lambda: dev.reduce_schur_dense_mfma_(Sq, gq, brq, beq, seq, q.lam, hq, Eq, Cq, wq,
                                                                      check=False)))
    for name, fn in legs:
        ts = []
        for rep_ in range(reps + 1):
            dev.reset_schur_dense_(Sq, Rq, hq, srq, q.lam, check=False)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            assert fn()
            ts.append(time.perf_counter() - t1)
        ms = float(np.median(ts[1:])) * 1e3
        small[name] = {"ms": ms, "landmarks_per_s": sq.neblocks / (ms * 1e-3), "useful_TFLOPs": useful_q / (ms * 1e-3) / 1e12}
        log(f"dense leg at {sq.neblocks} landmarks, {name}: {ms:.3f} ms ({sq.neblocks / (ms * 1e-3):.4g} landmarks/s)")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = cpu_share()[0]
        args = oracle.SchurArgs(sq, b_ws(q, threads), None)
        So, go = np.zeros(sq.nr * sq.nr), np.zeros(sq.nr)
        E_inv, z, W = np.zeros(len(q.E_data)), np.zeros(sq.ne), np.zeros(len(q.C_data))
        Rcm = np.ascontiguousarray(q.R.T.reshape(-1))
        t1 = time.perf_counter()
        oracle.reset_schur_dense(So, Rcm, q.scaling_r, q.lam)
        okc, _ = oracle.reduce_schur(False, args, So, go, q.b_r, q.b_e, q.scaling_e, q.lam, q.E_data, q.C_data, E_inv, z,
                                     W, nthreads=threads)
        cms = (time.perf_counter() - t1) * 1e3
        assert okc
        dev.reset_schur_dense_(Sq, Rq, hq, srq, q.lam, check=False)
        assert dev.reduce_schur_dense_(Sq, gq, brq, beq, seq, q.lam, hq, Eq, Cq, wq, check=False)
        A = Sq.cpu().numpy().reshape(sq.nr, sq.nr).T
        B = So.reshape(sq.nr, sq.nr).T
        iu = np.triu_indices(sq.nr)
        small["cpu_baseline"] = {"ms": cms, "landmarks_per_s": sq.neblocks / (cms * 1e-3), "threads": threads,
                                 "kind": "port", "sample": f"reset + reduce of the {sq.neblocks}-landmark instance"}
        small["parity_valu_vs_oracle"] = {"rel_S": float(np.linalg.norm(A[iu] - B[iu]) / np.linalg.norm(B[iu])),
                                          "rel_g": float(np.linalg.norm(gq.cpu().numpy() - go) / np.linalg.norm(go))}
        log(f"dense leg at {sq.neblocks} landmarks, CPU oracle ({threads} threads): {cms:.1f} ms; device S within "
            f"{small['parity_valu_vs_oracle']['rel_S']:.2e} of it (norm-wise)")
    out["owner_computes"] = small
    return out


def time_ms(torch, fn, reps=5, batch=1):
    """Median HIP-event time of fn() on the current stream, after one untimed call. batch > 1: each sample times
    that many calls back to back (per call), so that the Python wrapper's host work overlaps the device's."""
    fn()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(batch):
            fn()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) / batch)
    return float(np.median(times))


def assembly_leg(torch, dev, st, reps=10):
    """The SLAM workload's Schur Hessian assembled on the device from its row blocks (batched addblock!,
    src/schur_assembly.jl:140-246): one 2-row observation block per coupling ([landmark; camera], J 2 x 9) and a
    6-row prior per camera, R = the camera blocks; J, r ~ N(0, 1). HBM figure: J and r read once, E_data, C_data,
    R and b written once."""
    t0 = time.time()
    nrb, neb = int(st.nrblocks), int(st.neblocks)
    ep = np.asarray(st.eliminated_ptr) - 1
    er = np.asarray(st.entry_retained, dtype=np.int64)
    rd = np.asarray(st.r_dims, dtype=np.int64)
    ed = np.asarray(st.e_dims, dtype=np.int64)
    nobs = len(er)
    e_of = np.repeat(np.arange(neb, dtype=np.int64), np.diff(ep))
    rb_ptr = np.concatenate([1 + 2 * np.arange(nobs + 1, dtype=np.int64), 2 * nobs + 2 + np.arange(nrb, dtype=np.int64)])
    rb_blocks = np.concatenate([np.stack([nrb + e_of + 1, er], axis=1).reshape(-1), np.arange(1, nrb + 1, dtype=np.int64)])
    rb_rows = np.concatenate([np.full(nobs, 2, dtype=np.int64), rd])
    roff = np.concatenate([[1], 1 + np.cumsum(rd * rd)])[:-1]
    R_layout = (np.arange(1, nrb + 2), np.arange(1, nrb + 1), roff)
    h = dev.AssemblyHandle(st, (rb_ptr, rb_blocks), rb_rows, R_layout)
    nJ = int(np.sum(2 * (ed[e_of] + rd[er - 1])) + np.sum(rd * rd))
    nres = int(2 * nobs + np.sum(rd))
    gen = torch.Generator(device="cuda").manual_seed(7)
    J = torch.randn(nJ, dtype=torch.float64, device="cuda", generator=gen)
    r = torch.randn(nres, dtype=torch.float64, device="cuda", generator=gen)
    E = torch.empty(h.E_len, dtype=torch.float64, device="cuda")
    C = torch.empty(h.C_len, dtype=torch.float64, device="cuda")
    R = torch.empty(h.R_len, dtype=torch.float64, device="cuda")
    br = torch.empty(h.nr, dtype=torch.float64, device="cuda")
    be = torch.empty(int(ed.sum()), dtype=torch.float64, device="cuda")
    setup_s = time.time() - t0
    dev.assemble_hessian_(E, C, R, br, be, h, J, r)
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dev.assemble_hessian_(E, C, R, br, be, h, J, r)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    nbytes = 8.0 * (nJ + nres + h.E_len + h.C_len + h.R_len + h.nr + int(ed.sum()))
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"assembly (batched addblock!): {nobs} observation row blocks + {nrb} priors: {ms:.3f} ms, "
        f"{gbs:.0f} GB/s of {nbytes / 1e9:.3f} GB (plan {setup_s:.1f} s)")
    # dot(x, H, y) over the storage just filled: E_data, C_data, R read once, x and y read once
    n = h.nr + int(ed.sum())
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    dev.hessian_dot(x, h, y, E, C, R)
    dtimes = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dev.lib().bls_asm_dot_f64(h.h, dev._dp(x), dev._dp(y), dev._dp(E), dev._dp(C), dev._dp(R), dev._dp(out),
                                  dev._stream(None))
        e1.record()
        torch.cuda.synchronize()
        dtimes.append(e0.elapsed_time(e1))
    dms = float(np.median(dtimes))
    dbytes = 8.0 * (h.E_len + h.C_len + h.R_len + 2 * n)
    log(f"dot(x, H, y) on that storage: {dms:.4f} ms, {dbytes / (dms * 1e-3) / 1e9:.0f} GB/s")
    # the conditioning scale of a fresh fill: getdiag! (the diagonals of R and E_data) then prepare_scale! (d and sc
    # read, s written); and at a later damping, the drift test (d, sc, s read) with no recomputation
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    sc = torch.ones(n, dtype=torch.float64, device="cuda")
    c = dev.EliminationConditioning(n)
    diag_ = lambda dd: dev.hessian_diag_(dd, h, E, R)

    def fresh():
        c.markstale()
        dev.prepare_scale_(c, diag_, 0.0, sc)
    fms = time_ms(torch, fresh, reps)
    lam = 1e-3 * float(c.d.max())
    dev.prepare_scale_(c, diag_, lam, sc)
    later_ms = time_ms(torch, lambda: dev.prepare_scale_(c, diag_, lam, sc), reps)
    diag_ms = time_ms(torch, lambda: diag_(d), reps)
    assert c.nrescale() == 0
    fbytes = 8.0 * (h.nr + int(ed.sum()) + n + 3 * n)   # the diagonal entries read, d written; d, sc read, s written
    log(f"conditioning scale: getdiag! {diag_ms:.4f} ms; fresh fill (getdiag! + scale) {fms:.4f} ms, "
        f"{fbytes / (fms * 1e-3) / 1e9:.0f} GB/s; later damping (drift test) {later_ms:.4f} ms, "
        f"{24.0 * n / (later_ms * 1e-3) / 1e9:.0f} GB/s")
    return {"row_blocks": int(nobs + nrb), "ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "bytes": nbytes, "landmarks_per_s": neb / (ms * 1e-3),
            "dot": {"ms": dms, "achieved_GBs": dbytes / (dms * 1e-3) / 1e9, "bytes": dbytes},
            "conditioning": {"variables": n, "getdiag_ms": diag_ms, "fresh_fill_ms": fms, "later_damping_ms": later_ms,
                             "fresh_fill_GBs": fbytes / (fms * 1e-3) / 1e9,
                             "later_damping_GBs": 24.0 * n / (later_ms * 1e-3) / 1e9}}


def dense_leg(torch, dev, n, reps=5):
    """The dense reduced solve at the SLAM system's order (n = nr): Cholesky of a copy of S (the :cholesky branch of
    solve_reduced_dense!) and the two triangular solves, on a synthetic SPD S = A A'/n + I (A ~ N(0, 1)). FLOP figure:
    n^3 / 3 for the factorization (its trailing updates run on the FP64 matrix instruction)."""
    gen = torch.Generator(device="cuda").manual_seed(11)
    A = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=gen)
    S = (A @ A.T / n + torch.eye(n, dtype=torch.float64, device="cuda")).contiguous()
    del A
    g = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    ch = dev.DenseCholesky(n)
    ok, _ = ch.factor(S)
    assert ok
    st = dev._stream(None)
    fms = time_ms(torch, lambda: dev.lib().bls_chol_factor_f64_async(ch.h, dev._dp(S), st), reps)
    sms = time_ms(torch, lambda: ch.solve(g, x), reps)
    rel = float(torch.linalg.norm(S @ x - g) / torch.linalg.norm(g))
    tf = n ** 3 / 3.0 / (fms * 1e-3) / 1e12
    log(f"dense reduced solve, n = {n}: Cholesky {fms:.3f} ms ({tf:.1f} TFLOP/s, {100 * tf / FP64_MFMA_PEAK_TF:.1f}% of "
        f"the FP64 matrix peak), solve {sms:.3f} ms, ||S x - g|| / ||g|| = {rel:.1e}")
    # the :ldl fallback (Bunch-Kaufman) on an indefinite S of the same order: S - 1.5 I has negative eigenvalues
    S.sub_(1.5 * torch.eye(n, dtype=torch.float64, device="cuda"))
    ldl = dev.DenseLDL(n)
    t0 = time.time()
    ok, neg = ldl.factor(S)
    lfms = (time.time() - t0) * 1e3   # bls_ldl_factor_f64 synchronizes: host wall time of the whole factorization
    assert ok and neg > 0
    lsms = time_ms(torch, lambda: ldl.solve(g, x), reps)
    lrel = float(torch.linalg.norm(S @ x - g) / torch.linalg.norm(g))
    log(f"dense reduced solve, n = {n}: Bunch-Kaufman (:ldl fallback) {lfms:.1f} ms ({neg} negative eigenvalues), "
        f"solve {lsms:.3f} ms, ||S x - g|| / ||g|| = {lrel:.1e}")
    return {"n": n, "factor_ms": fms, "solve_ms": sms, "factor_TFLOPs": tf, "frac_fp64_mfma": tf / FP64_MFMA_PEAK_TF,
            "relres": rel, "ldl": {"factor_ms": lfms, "solve_ms": lsms, "negative_eigenvalues": neg, "relres": lrel}}


def mixed_leg(torch, dev, p, h, S64, csc, dmap, steps, warmup):
    """The same step with the blocks E, C, R stored in Float32 (the mixed entry points: Float64 landmark
    factorisation, S and g_S; the reduced system of the Float32 blocks to Float64 rounding). Reports its rate, its
    kernel times and how far its S is from the Float64 step's (the rounding of the inputs to Float32)."""
    st = p.structure
    ws = dev.SchurReductionWorkspace(h, "mixed")
    T32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    T64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, Rd = T32(p.E_data), T32(p.C_data), T32(p.R_data)
    b_r, b_e, sr, se = T64(p.b_r), T64(p.b_e), T64(p.scaling_r), T64(p.scaling_e)
    S = torch.empty_like(S64)
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    ref = S64.clone()

    def step():
        dev.reset_schur_sparse_(S, Rd, h, sr, p.lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, sync=False, check=False)
        dev.copyto_(csc, S, dmap)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    assert int(ws.d_failed.item()) == 0
    h.set_timing(True)
    kt = np.zeros(4)
    for _ in range(steps):
        dev.reset_schur_sparse_(S, Rd, h, sr, p.lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, check=False)
        kt += np.array(h.kernel_times())
    h.set_timing(False)
    kt /= steps
    k = np.diff(st.eliminated_ptr)
    # phase 2 reads E_inv (Float64, 72 B), z (24 B) and C (Float32, 72 B per coupling), reads and writes S
    p2_bytes = float(np.sum(96 + 72 * k) + 16 * p.pattern.nnz + 16 * st.nr)
    rel = float((S - ref).norm() / ref.norm())
    log(f"mixed (Float32 blocks, Float64 system): {ms:.4f} ms/step, {st.neblocks / (ms * 1e-3):.4g} landmarks/s; "
        f"phase1 {kt[1]:.4f} ms, phase2 {kt[2]:.4f} ms ({p2_bytes / (kt[2] * 1e-3) / 1e9:.0f} GB/s); "
        f"||S_mixed - S_f64|| / ||S_f64|| = {rel:.3g}")
    return {"ms_per_step": ms, "landmarks_per_s": st.neblocks / (ms * 1e-3),
            "kernel_ms": {"reset": kt[0], "phase1": kt[1], "phase2": kt[2], "phase2_merge": kt[3]},
            "phase2_GBs": p2_bytes / (kt[2] * 1e-3) / 1e9, "rel_diff_S_vs_f64_inputs": rel,
            "dtype": "E, C, R Float32; E_inv, S, g Float64"}


def fill_leg(torch, dev, p, raw, r, reps=10):
    """The tables filled on the device from the workload's per-observation blocks (block / residual_block of every
    row block, bls_jfill_f64): the generator's 2 x d camera and 2 x 3 landmark blocks laid out row block after row
    block, written into fresh tables and compared bit for bit with the tables the generator built. HBM figure: the
    values and residual rows read once and written once (16 B per value)."""
    jst, st = p.jst, p.structure
    fill = dev.JacobianFill(jst)
    d = st.r_dims[p.cam_of_obs - 1].astype(np.int64)
    n_k = 2 * d + 6
    off = np.concatenate([[0], np.cumsum(n_k)[:-1]])
    vals = np.empty(int(n_k.sum()))
    for c in range(int(d.max())):
        m = c < d
        for i in range(2):
            vals[off[m] + 2 * c + i] = p.Jc[m, i, c]
    for c in range(3):
        for i in range(2):
            vals[off + 2 * d + 2 * c + i] = p.Jl[:, i, c]
    assert len(vals) == fill.nvals
    dv = torch.as_tensor(vals).cuda()
    dres = torch.as_tensor(np.ascontiguousarray(p.rr.reshape(-1))).cuda()
    del vals
    out = torch.zeros_like(raw)
    rout = torch.zeros_like(r)
    dev.fill_jacobian_(out, rout, fill, dv, dres)
    torch.cuda.synchronize()
    same = bool(torch.equal(out, raw) and torch.equal(rout, r))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(reps):
        e0.record()
        dev.fill_jacobian_(out, rout, fill, dv, dres)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = float(np.median(times))
    nbytes = 16.0 * (fill.nvals + fill.nres)
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"table fill: {ms:.3f} ms for {len(p.cam_of_obs)} row blocks ({fill.nvals} values), {gbs:.0f} GB/s "
        f"({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM); equal to the generator's tables: {same}")
    if not same:
        raise SystemExit("table fill differs from the generator's tables")
    del out, rout, dv, dres
    return {"row_blocks": int(len(p.cam_of_obs)), "values": int(fill.nvals), "ms": ms, "achieved_GBs": gbs,
            "frac": gbs / HBM_PEAK_GBS, "bytes": nbytes, "bitwise_equal_to_generator_tables": same}


def qr_leg(torch, dev, b, scale, reps=10):
    """reduce_qr! on the Jacobian workload: 500 cameras (d 6 or 9, 50/50), 500000 x scale landmarks (d 3), track
    length U{2..30}, 2 x d residual blocks N(0,1), lambda 1e-4, FP64. Time of the table kernel (HIP events) and its
    HBM figure: per table 2 * 8 M C bytes (raw in, factored out) + 16 M (r in, fres out)."""
    t0 = time.time()
    p = b.generators.config4(scale=scale)
    jst, st = p.jst, p.structure
    h = dev.QRHandle(jst)
    ws = dev.QRReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    raw, r, sc = T(p.data), T(p.r), T(p.scaling_e)
    log(f"QR leg setup: {jst.ntables} tables ({st.neblocks} landmarks), {st.nrblocks} cameras, raw "
        f"{jst.raw_length() * 8 / 1e9:.2f} GB ({time.time() - t0:.1f}s)")
    lam = 1e-4
    assert dev.reduce_qr_(ws, h, raw, r, sc, lam)
    h.set_timing(True)
    times = []
    for _ in range(reps):
        assert dev.reduce_qr_(ws, h, raw, r, sc, lam, check=False)
        times.append(h.kernel_time())
    h.set_timing(False)
    ms = float(np.median(times))
    M = jst.table_nrows.astype(np.float64)
    Cc = (jst.table_ncols - 1).astype(np.float64)
    nbytes = float(np.sum(16 * M * Cc + 16 * M))
    gbs = nbytes / (ms * 1e-3) / 1e9
    log(f"QR leg: {ms:.3f} ms, {st.neblocks / (ms * 1e-3):.4g} landmarks/s, {gbs:.0f} GB/s "
        f"({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM)")
    fill = fill_leg(torch, dev, p, raw, r)
    # SchurBlockJacobian{Float32}: Float32 tables, Float64 residual
    F64 = ws.factored.clone()
    raw64 = raw
    ws32 = dev.QRReductionWorkspace(h, dtype=torch.float32)
    raw32 = torch.as_tensor(np.ascontiguousarray(p.data, dtype=np.float32)).cuda()
    assert dev.reduce_qr_(ws32, h, raw32, r, sc, lam)
    h.set_timing(True)
    times = []
    for _ in range(reps):
        assert dev.reduce_qr_(ws32, h, raw32, r, sc, lam, check=False)
        times.append(h.kernel_time())
    h.set_timing(False)
    ms32 = float(np.median(times))
    nbytes32 = float(np.sum(8 * M * Cc + 16 * M))
    gbs32 = nbytes32 / (ms32 * 1e-3) / 1e9
    rel32 = float((ws32.factored.double() - F64).norm() / F64.norm())
    log(f"QR leg, Float32 tables: {ms32:.3f} ms, {st.neblocks / (ms32 * 1e-3):.4g} landmarks/s, {gbs32:.0f} GB/s; "
        f"||F32 - F64|| / ||F64|| = {rel32:.3g}")
    f32 = {"ms": ms32, "landmarks_per_s": st.neblocks / (ms32 * 1e-3), "achieved_GBs": gbs32,
           "frac": gbs32 / HBM_PEAK_GBS, "rel_diff_vs_f64": rel32}
    # the reduced system on the device (reduced_normal_sparse!, plus lambda on the retained diagonal) and block-Jacobi
    # PCG to 1e-8: time to 1e-8 from the Jacobian, with Float64 tables and with Float32 tables
    pat = b.structure.build_schur_pattern(st)
    h.set_sparse(pat)
    pcg = dev.PCGHandle(st, pat)
    diag = torch.as_tensor(pat.diag_inds - 1).cuda()
    solve = {}
    for name, wsx, raw_x, qms in (("f64", ws, raw64, ms), ("f32", ws32, raw32, ms32)):
        S = torch.empty(pat.nnz, dtype=wsx.factored.dtype, device="cuda")
        gS = torch.empty(st.nr, dtype=torch.float64, device="cuda")
        x = torch.zeros(st.nr, dtype=torch.float64, device="cuda")
        t_n, t_p = [], []
        for rep_ in range(4):
            assert dev.reduce_qr_(wsx, h, raw_x, r, sc, lam, check=False)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dev.reduced_normal_sparse_(S, gS, h, wsx)
            S64 = S if S.dtype == torch.float64 else S.double()
            S64[diag] += lam
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            x.zero_()
            ok, iters, rel = pcg.solve(S64, gS, x, tol=1e-8, maxit=5000)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep_:
                t_n.append(t1 - t0)
                t_p.append(t2 - t1)
        nm, pm = float(np.median(t_n)) * 1e3, float(np.median(t_p)) * 1e3
        solve[name] = {"reduced_normal_ms": nm, "pcg_ms": pm, "pcg_iters": int(iters.max()), "relres": float(rel.max()),
                       "converged": bool(ok), "time_to_1e-8_ms": qms + nm + pm if ok else None}
        log(f"QR leg, {name} tables: reduced_normal_sparse! + lambda I {nm:.3f} ms, PCG {pm:.3f} ms ({int(iters.max())} "
            f"iterations), time to 1e-8 from the Jacobian {qms + nm + pm:.3f} ms")
    # products with the reduced Jacobian on the Float64 tables: each reads the retained columns of every table's
    # reduced rows once (8 B per value) plus its vectors
    assert dev.reduce_qr_(ws, h, raw64, r, sc, lam, check=False)
    nred = int(jst.reduced_nrows())
    ncols = np.bincount(np.asarray(jst.entry_table) - 1, weights=np.asarray(st.r_dims)[np.asarray(jst.entry_retained) - 1],
                        minlength=jst.ntables)
    xbytes = float(np.sum((jst.table_nrows - jst.table_top) * ncols) * 8)
    rng = np.random.default_rng(5)
    vv = T(rng.standard_normal(st.nr))
    ww = T(rng.standard_normal(nred))
    yv = torch.empty(nred, dtype=torch.float64, device="cuda")
    gv = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    Dv = torch.empty(int((np.asarray(st.r_dims) ** 2).sum()), dtype=torch.float64, device="cuda")
    prods = {}
    jcp, jrv = dev.build_reduced_jacobian_pattern(h)
    dev.check_reduced_jacobian_pattern(h, nred, st.nr, jcp, jrv)
    jnz = torch.empty(len(jrv), dtype=torch.float64, device="cuda")
    jrt = torch.empty(nred, dtype=torch.float64, device="cuda")
    for pname, fn, vec_bytes in (("reduced_mul", lambda: dev.reduced_mul_(yv, h, ws.factored, vv), 8 * nred),
                                 ("reduced_jacobian_sparse", lambda: dev.reduced_jacobian_(jnz, jrt, h, ws, sparse=True),
                                  xbytes + 16 * nred),
                                 ("reduced_tmul", lambda: dev.reduced_tmul_(gv, h, ws.factored, ww), 8 * nred),
                                 ("reduced_diagonal_blocks", lambda: dev.reduced_diagonal_blocks_(Dv, h, ws.factored),
                                  8 * Dv.numel())):
        fn()
        times = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        pms = float(np.median(times))
        gb = (xbytes + vec_bytes) / (pms * 1e-3) / 1e9
        prods[pname] = {"ms": pms, "achieved_GBs": gb, "frac": gb / HBM_PEAK_GBS}
        log(f"QR leg, {pname}: {pms:.3f} ms, {gb:.0f} GB/s of {(xbytes + vec_bytes) / 1e9:.2f} GB")
    solve["products_f64"] = prods
    # the QR backsubstitute! (x_e = U^-1 (q - B x_r) per eliminated table): the first d_e rows of every column of
    # each eliminated table, and q, read once; x_r gathered, x_e written
    xr = T(rng.standard_normal(st.nr))
    xe = torch.empty(st.ne, dtype=torch.float64, device="cuda")
    ed_ = np.asarray(st.e_dims, dtype=np.int64)
    qbytes = 8.0 * float(np.sum(ed_ * (np.asarray(jst.table_ncols[:st.neblocks]) - 1) + ed_) + st.nr + st.ne)
    qbms = time_ms(torch, lambda: dev.qr_backsubstitute_(xe, xr, h, ws), batch=10)
    log(f"QR leg, backsubstitute!: {qbms:.3f} ms, {qbytes / (qbms * 1e-3) / 1e9:.0f} GB/s of {qbytes / 1e9:.3f} GB")
    solve["backsubstitute_f64"] = {"ms": qbms, "achieved_GBs": qbytes / (qbms * 1e-3) / 1e9, "bytes": qbytes}
    # quantities of the full problem from the raw Float64 tables: each reads every table once (8 B per value) plus
    # its vectors
    nv = st.nr + st.ne
    rawbytes = float(jst.raw_length() * 8)
    gq = torch.empty(nv, dtype=torch.float64, device="cuda")
    vq = T(rng.standard_normal(nv))
    full = {}
    for pname, fn, extra in (("jtr", lambda: dev.jtr_(gq, h, raw64, r), 8 * (jst.residual_length() + nv)),
                             ("jtj_diag", lambda: dev.jtj_diag_(gq, h, raw64), 8 * nv),
                             ("jtj_quadform", lambda: dev.jtj_quadform(h, raw64, vq), 8 * nv)):
        pms = time_ms(torch, fn, batch=5)
        gb = (rawbytes + extra) / (pms * 1e-3) / 1e9
        full[pname] = {"ms": pms, "achieved_GBs": gb, "frac": gb / HBM_PEAK_GBS}
        log(f"QR leg, {pname} (raw tables): {pms:.3f} ms, {gb:.0f} GB/s of {(rawbytes + extra) / 1e9:.2f} GB")
    solve["full_problem_f64"] = full
    return {"landmark_scale": scale, "f32": f32, "fill": fill, "solve": solve, "tables": int(jst.ntables),
            "landmarks": int(st.neblocks),
            "cameras": int(st.nrblocks), "raw_GB": jst.raw_length() * 8 / 1e9, "ms": ms,
            "landmarks_per_s": st.neblocks / (ms * 1e-3),
            "roofline": {"bound": "hbm", "kernel": "k_qr_tables", "achieved": gbs, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}}


def batched_pcg_leg(torch, dev, b, scale, reps=3, check_parity=False):
    """64 independent problems (camera counts U{50..2000}, landmark counts U{10k..300k} x scale, SLAM tracks,
    lambda 1e-4) as one block-diagonal structure: one reduction (one launch per kernel), then block-Jacobi PCG on all
    64 reduced systems at once to ||S x - g|| <= 1e-8 ||g|| per problem (FP64 true residual). Reports the
    device time of each part and their sum, the time to 1e-8."""
    t0 = time.time()
    p = b.generators.config5_union(scale=scale)
    st = p.structure
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h)
    pcg = dev.PCGHandle(st, p.pattern, p.prob_block_ptr)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, b_r, b_e = T(p.E_data), T(p.C_data), T(p.b_r), T(p.b_e)
    sr, se, Rd = T(p.scaling_r), T(p.scaling_e), T(p.R_data)
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    x = torch.zeros(st.nr, dtype=torch.float64, device="cuda")
    log(f"batched leg setup: {p.nprob} problems, {st.nrblocks} cameras, {st.neblocks} landmarks, S nnz {p.pattern.nnz} "
        f"({time.time() - t0:.1f}s)")
    t0 = time.time()
    dev.reset_schur_sparse_(S, Rd, h, sr, p.lam)
    cands = [(384, 96), (1536, 96), (2304, 64), (2304, 96)]   # each shape rebuilds the work list of 9.85 M landmarks
    tuned = h.autotune(cands, S, g, b_r, b_e, se, p.lam, E, C, ws)
    log(f"batched leg autotune: part {tuned[0]}, lanes {tuned[1]} ({tuned[2]:.3f} ms reduce) over {len(cands)} shapes "
        f"({time.time() - t0:.1f}s)")

    def reduce():
        dev.reset_schur_sparse_(S, Rd, h, sr, p.lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, sync=False, check=False)

    reduce()
    x.zero_()
    pcg.solve(S, g, x, tol=1e-8)
    torch.cuda.synchronize()
    t_red, t_pcg = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        reduce()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x.zero_()
        ok, iters, rel = pcg.solve(S, g, x, tol=1e-8)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        assert ok and int(ws.d_failed.item()) == 0
        t_red.append(t1 - t0)
        t_pcg.append(t2 - t1)
    red, sol = float(np.median(t_red)) * 1e3, float(np.median(t_pcg)) * 1e3
    log(f"batched leg: reduce {red:.3f} ms ({st.neblocks / red * 1e3:.4g} landmarks/s), PCG {sol:.3f} ms "
        f"(iterations max {int(iters.max())}, mean {iters.mean():.1f}), time to 1e-8 {red + sol:.3f} ms")
    learned = None
    P = b.precond
    if os.path.exists(P.WEIGHTS):
        # the learned correction: features, MLP, corrected blocks and PCG, all on the device (falls back to plain
        # block-Jacobi if its solve fails)
        model = P.MLP.load()
        P.solve(pcg, S, g, x, p.lam, model=model)
        torch.cuda.synchronize()
        t_l = []
        for _ in range(reps):
            t0 = time.perf_counter()
            okl, itl, rell, used = P.solve(pcg, S, g, x, p.lam, model=model)
            torch.cuda.synchronize()
            t_l.append(time.perf_counter() - t0)
        lms = float(np.median(t_l)) * 1e3
        learned = {"pcg_ms": lms, "time_to_1e-8_ms": red + lms, "pcg_iters_max": int(itl.max()),
                   "pcg_iters_mean": float(itl.mean()), "relres_max": float(rell.max()), "used_learned": bool(used),
                   "output_scale": model.alpha}
        log(f"batched leg, learned correction (output scale {model.alpha}): PCG {lms:.3f} ms (iterations max "
            f"{int(itl.max())}, mean {itl.mean():.1f}), time to 1e-8 {red + lms:.3f} ms")
    mixed = batched_mixed(torch, dev, b, p, h, pcg, S, g, reps, f64=dict(E=E, C=C, R=Rd, b_r=b_r, b_e=b_e, sr=sr, se=se))
    parity = None
    if check_parity:
        # the union's reduction at full size against the oracle's (the device's S and g of the last timed reduce)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        t0 = time.time()
        threads = cpu_share()[0]
        args = oracle.SchurArgs(st, b_ws(p, threads), p.pattern)
        S_ref, g_ref = np.zeros(p.pattern.nnz), np.zeros(st.nr)
        E_inv, z, W = np.zeros(len(p.E_data)), np.zeros(st.ne), np.zeros(len(p.C_data))
        oracle.reset_schur_sparse(S_ref, p.R_data, p.r_to_s, p.pattern, p.scaling_r, p.lam)
        ok, _ = oracle.reduce_schur(True, args, S_ref, g_ref, p.b_r, p.b_e, p.scaling_e, p.lam, p.E_data, p.C_data,
                                    E_inv, z, W, nthreads=threads)
        assert ok
        parity = parity_record(st, p.pattern, S.cpu().numpy(), g.cpu().numpy(), S_ref, g_ref)
        log(f"batched leg parity against the oracle: S max block rel {parity['max_block_rel']:.3g}, g max block rel "
            f"{parity['g_max_block_rel']:.3g} ({time.time() - t0:.1f}s)")
    return {"problems": int(p.nprob), "landmark_scale": scale, "cameras": int(st.nrblocks),
            "landmarks": int(st.neblocks), "S_nnz": int(p.pattern.nnz), "reduce_ms": red,
            "landmarks_per_s": st.neblocks / (red * 1e-3), "pcg_ms": sol, "time_to_1e-8_ms": red + sol,
            "pcg_iters_max": int(iters.max()), "pcg_iters_mean": float(iters.mean()),
            "relres_max": float(rel.max()), "preconditioner": "block-Jacobi (FP64)", "learned": learned,
            "tuning": {"part": tuned[0], "lanes": tuned[1]},
            "mixed": mixed, "parity": parity}


def batched_mixed(torch, dev, b, p, h, pcg, S64, g64, reps, f64=None):
    """The 64 problems with their blocks E, C, R stored in Float32: the mixed reduction (Float64 factorisation and
    reduced system of the Float32 blocks), then the same PCG (plain and learned) to 1e-8 on that system. Also the
    residual of its solution against the Float64-stored problem's system, ||S64 x - g64|| / ||g64|| per problem:
    the cost of the Float32 inputs."""
    st = p.structure
    ws = dev.SchurReductionWorkspace(h, "mixed")
    T32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    T64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, Rd = T32(p.E_data), T32(p.C_data), T32(p.R_data)
    b_r, b_e, sr, se = T64(p.b_r), T64(p.b_e), T64(p.scaling_r), T64(p.scaling_e)
    S = torch.empty_like(S64)
    g = torch.empty_like(g64)
    x = torch.zeros_like(g64)

    def reduce():
        dev.reset_schur_sparse_(S, Rd, h, sr, p.lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, sync=False, check=False)

    reduce()
    x.zero_()
    pcg.solve(S, g, x, tol=1e-8)
    torch.cuda.synchronize()
    t_red, t_pcg = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        reduce()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        x.zero_()
        ok, iters, rel = pcg.solve(S, g, x, tol=1e-8)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        assert ok and int(ws.d_failed.item()) == 0
        t_red.append(t1 - t0)
        t_pcg.append(t2 - t1)
    red, sol = float(np.median(t_red)) * 1e3, float(np.median(t_pcg)) * 1e3
    # residual against the Float64-stored problem's reduced system (a check, untimed)
    cross = float(pcg.relres(S64, g64, x).max())
    own = float(pcg.relres(S, g, x).max())
    out = {"reduce_ms": red, "landmarks_per_s": st.neblocks / (red * 1e-3), "pcg_ms": sol,
           "time_to_1e-8_ms": red + sol, "pcg_iters_max": int(iters.max()), "pcg_iters_mean": float(iters.mean()),
           "relres_max": float(rel.max()), "relres_recomputed_max": own, "relres_vs_f64_inputs_max": cross,
           "dtype": "E, C, R Float32; E_inv, S, g, PCG Float64"}
    log(f"batched leg, mixed: reduce {red:.3f} ms ({st.neblocks / red * 1e3:.4g} landmarks/s), PCG {sol:.3f} ms "
        f"(iterations max {int(iters.max())}), time to 1e-8 {red + sol:.3f} ms; residual against the Float64-stored "
        f"system {cross:.3g}")
    P = b.precond
    if os.path.exists(P.WEIGHTS):
        model = P.MLP.load()
        P.solve(pcg, S, g, x, p.lam, model=model)
        torch.cuda.synchronize()
        t_l = []
        for _ in range(reps):
            t0 = time.perf_counter()
            okl, itl, rell, used = P.solve(pcg, S, g, x, p.lam, model=model)
            torch.cuda.synchronize()
            t_l.append(time.perf_counter() - t0)
        lms = float(np.median(t_l)) * 1e3
        out["learned"] = {"pcg_ms": lms, "time_to_1e-8_ms": red + lms, "pcg_iters_max": int(itl.max()),
                          "relres_max": float(rell.max()), "used_learned": bool(used), "output_scale": model.alpha}
        log(f"batched leg, mixed + learned: time to 1e-8 {red + lms:.3f} ms")
    if f64 is not None:
        out["refined"] = refined_leg(torch, dev, b, p, h, pcg, S, g, S64, g64, reduce, f64, reps)
    return out


def refined_leg(torch, dev, b, p, h, pcg, S_m, g_m, S64, g64, reduce, f64, reps):
    """The mixed reduction followed by Float64 iterative refinement to ||g - S x|| <= 1e-8 ||g|| against the
    Float64 blocks' reduced system: the Float64 factorisation of the landmark blocks, R + damping reset into the
    reduced pattern, then {PCG on the mixed system to 1e-4; x += d; the residual from the Float64 blocks} until
    met. The whole of it is timed; the formed Float64 system is used only to check the result (untimed)."""
    R = b.refine
    st = p.structure
    SR = torch.empty_like(S64)
    rd = R.RefinementData(h, f64["E"], f64["C"], f64["b_r"], f64["b_e"], f64["se"], SR, p.lam)
    x = torch.zeros_like(g64)
    P = b.precond
    model = P.MLP.load() if os.path.exists(P.WEIGHTS) else None

    def run(learned):
        reduce()
        dev.reset_schur_sparse_(SR, f64["R"], h, f64["sr"], p.lam, check=False)
        assert R.factor(rd, check=False)
        blocks = None
        if learned and model is not None and model.alpha != 0.0:
            blocks = P.group_blocks(pcg, S_m, model.omega(P.features(pcg, S_m, p.lam)), model.group)
        return R.refined_solve(pcg, S_m, g_m, x, rd, tol=1e-8, factored=True, group_blocks=blocks)

    res = {}
    for name, learned in (("plain", False), ("learned", True)):
        run(learned)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ok, steps, iters, rel = run(learned)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ms = float(np.median(ts)) * 1e3
        true_rel = float(pcg.relres(S64, g64, x).max())
        res[name] = {"time_to_1e-8_ms": ms, "converged": bool(ok), "steps": int(steps),
                     "pcg_iters_max": int(iters.max()), "relres_vs_f64_system_max": true_rel}
        log(f"batched leg, mixed + Float64 refinement ({name}): time to 1e-8 of the Float64 system {ms:.3f} ms, "
            f"{steps} steps, {int(iters.max())} PCG iterations, ||g - S x|| / ||g|| = {true_rel:.2e}")
    return res


def pcie_inclusive(torch, dev, p, h, ws, S, g, csc, dmap, steps):
    """The same step with the inputs in pinned host memory: copy E, C, b_r, b_e, R in, reduce, copy S, g and the
    CSC values out (the cost a caller holding host arrays pays; not the metric's value)."""
    st = p.structure
    pin = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).pin_memory()
    hE, hC, hbr, hbe, hR = pin(p.E_data), pin(p.C_data), pin(p.b_r), pin(p.b_e), pin(p.R_data)
    oS = torch.empty(S.numel(), dtype=torch.float64).pin_memory()
    og = torch.empty(g.numel(), dtype=torch.float64).pin_memory()
    ocsc = torch.empty(csc.numel(), dtype=torch.float64).pin_memory()
    dE, dC, dbr, dbe, dR = (torch.empty(x.numel(), dtype=torch.float64, device="cuda") for x in (hE, hC, hbr, hbe, hR))
    sr = torch.as_tensor(np.ascontiguousarray(p.scaling_r)).cuda()
    se = torch.as_tensor(np.ascontiguousarray(p.scaling_e)).cuda()

    def run():
        for d, x in ((dE, hE), (dC, hC), (dbr, hbr), (dbe, hbe), (dR, hR)):
            d.copy_(x, non_blocking=True)
        dev.reset_schur_sparse_(S, dR, h, sr, p.lam, check=False)
        dev.reduce_schur_sparse_(S, g, dbr, dbe, se, p.lam, h, dE, dC, ws, sync=False, check=False)
        dev.copyto_(csc, S, dmap)
        oS.copy_(S, non_blocking=True)
        og.copy_(g, non_blocking=True)
        ocsc.copy_(csc, non_blocking=True)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    moved = sum(x.numel() for x in (hE, hC, hbr, hbe, hR, oS, og, ocsc)) * 8
    return {"value": st.neblocks / (ms * 1e-3), "unit": "landmarks/s", "ms_per_step": ms,
            "host_bytes_per_step": moved}


def cpu_share():
    """Cores this process may use: its affinity set, bounded by a cgroup CPU quota when one is set, and by the 16
    cores that a one-GPU job is given on the GPU hosts. Returns (threads, description)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    share = min(aff, quota) if quota else aff
    threads = min(share, 16)
    return threads, {"nproc": os.cpu_count(), "affinity": aff, "cgroup_quota_cores": quota, "threads_all": threads}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_oracle(oracle, p, args, threads, budget_s, max_runs=7):
    """Median wall time of reset + reduce by the oracle with `threads` OpenMP threads; also returns S and g."""
    st = p.structure
    S = np.zeros(p.pattern.nnz)
    g = np.zeros(st.nr)
    E_inv = np.zeros(len(p.E_data))
    z = np.zeros(st.ne)
    W = np.zeros(len(p.C_data))
    times = []
    t_start = time.time()
    while len(times) < max_runs and (time.time() - t_start < budget_s or len(times) < 2):
        t0 = time.perf_counter()
        oracle.reset_schur_sparse(S, p.R_data, p.r_to_s, p.pattern, p.scaling_r, p.lam)
        ok, _ = oracle.reduce_schur(True, args, S, g, p.b_r, p.b_e, p.scaling_e, p.lam, p.E_data, p.C_data, E_inv, z, W,
                                    nthreads=threads)
        times.append(time.perf_counter() - t0)
        assert ok
    return float(np.median(times)), len(times), S, g


def cpu_baseline(p, threads):
    """The oracle (C restatement of the reference's threaded reduction, src/threading.jl:24-39 and
    ext/PolyesterExt.jl:12-23: phase 1 over nchunks chunks per run, phase 2 over the work-balanced column chunks)
    on the host cores: reset + reduce of the same instance, built -O3 -march=native on this host, timed with every
    core of the process's share and with one thread. Returns the baseline record and the oracle's S and g (all
    thread counts give the same bits)."""
    import shutil
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    st = p.structure
    share, cores = cpu_share()
    threads = threads or share
    cores["threads_all"] = threads
    tmp = tempfile.mkdtemp(prefix="blsorc_")
    build = oracle.use_native(tmp)
    shutil.rmtree(tmp, ignore_errors=True)   # the library stays mapped
    if build is None:
        build = "oracle/liborc.so (gcc -O2; the -O3 -march=native build failed on this host)"
    # chunked like the reference's threaded path (nchunks = threads), results identical to serial
    ws = b_ws(p, threads)
    args = oracle.SchurArgs(st, ws, p.pattern)
    t_all, n_all, S, g = time_oracle(oracle, p, args, threads, 15.0)
    ws1 = b_ws(p, 1)
    args1 = oracle.SchurArgs(st, ws1, p.pattern)
    t_one, n_one, S1, g1 = time_oracle(oracle, p, args1, 1, 20.0, max_runs=5)
    assert np.array_equal(S, S1) and np.array_equal(g, g1)
    log(f"CPU baseline ({cpu_model()}, {threads} threads of nproc {cores['nproc']}): {st.neblocks / t_all:.4g} "
        f"landmarks/s; 1 thread {st.neblocks / t_one:.4g} landmarks/s")
    rec = {"value": st.neblocks / t_all, "unit": "landmarks/s", "cores": threads, "kind": "port",
           "value_all": st.neblocks / t_all, "value_1thread": st.neblocks / t_one, "ms_all": t_all * 1e3,
           "ms_1thread": t_one * 1e3, "cpu_model": cpu_model(), **cores, "build": build,
           "sample": f"full workload (reset + reduce of {st.neblocks} landmarks), median of {n_all} runs at "
                     f"{threads} threads and of {n_one} runs at 1 thread; the C oracle with OpenMP over the "
                     f"reference's chunks"}
    return rec, S, g


def b_ws(p, threads):
    import blsload
    b = blsload.load()
    st = p.structure
    st2 = b.structure.SchurStructure(st.r_dims, st.e_dims, (st.eliminated_ptr - 1, st.entry_retained), nchunks=threads)
    st.chunks = st2.chunks
    return b.structure.SchurReductionWorkspaceIndex(st, threads)


if __name__ == "__main__":
    main()
