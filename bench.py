"""Benchmark: 3D variable-coefficient Poisson on 128^3 cells (2,097,152
unknowns, 14,581,760 nonzeros), k = exp(g), g ~ U[-ln 10, ln 10], Dirichlet
on z = 0, b ~ U[-1, 1]; smoothed-aggregation AMG (three-phase aggregates,
Jacobi 2/2, FP64 direct coarsest solve) preconditioned CG to a relative
residual of 1e-10.

A step is one solve from x = 0 with the matrix, hierarchy and right-hand side
resident in HBM.  ``value`` is the FP32-inner / FP64-refined time-to-1e-10 in
milliseconds; the plain FP64 time, their ratio, each solution's error against
the CPU FP64 oracle's solution of the same system and the solve-level
bandwidth are reported beside it.  ``roofline`` is the FP64 CSR SpMV y = A x
of the fine operator (row blocks, int32 row pointers and columns: the layout
whose bytes nnz (s + 4) + n (2 s + 4) + 4 define the metric), timed with HIP
events on the stream it is launched on; ``spmv_device_format`` is the same
product in the format the solver uses for this stencil (symmetric diagonal
storage).  ``cpu_baseline`` is the CPU FP64 oracle solving the same system
on the host's cores (up to 16 threads), its one-thread time beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the SpMV kernels, from two rocprofv3 counter passes
# over `bench.py --spmv-only --spmv-format F`: FETCH_SIZE x 2.0 + WRITE_SIZE x
# 1.0 KiB, the factors that make the counters read 512 MiB each for a 512 MiB
# device copy.  Keyed by the device format measured (128^3 operator).
SPMV_FP64_TRAFFIC_BYTES = {"dia": round((42329.487 * 2.0 + 16384.0) * 1024),
                           "csr_int32": round((99720.64 * 2.0 + 16384.0) * 1024)}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int, s: int, storage: dict) -> int:
    """Bytes one y = A x must move in the matrix's device format, x read once
    and y written once (s = bytes per value)."""
    if storage["format"] == "dia":
        # stored diagonals, no indices (the lower half of a symmetric matrix is re-read from the upper)
        return storage["stored_entries"] * s + 2 * n * s
    if storage["format"] == "sell64":
        # padded slots: value + int16 column offset; one slice pointer per slice
        return storage["stored_entries"] * (s + 2) + (storage["units"] + 1) * 4 + 2 * n * s
    # CSR with int32 columns: the metric's model nnz (s + 4) + n (2 s + 4) + 4
    # (value + column per nonzero, row pointers, x read once, y written once)
    return nnz * (s + 4) + n * (2 * s + 4) + 4


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--theta", type=float, default=0.25)
    ap.add_argument("--omega", type=float, default=0.8)
    ap.add_argument("--inner-tol", type=float, default=1e-5,
                    help="inner tolerance of restarted refinement (--reliable-interval -1)")
    ap.add_argument("--reliable-delta", type=float, default=1e-2,
                    help="FP32-refined CG: replace the residual when it fell by this factor (0 = library default, "
                         "1e-1); 1e-2 measured best on this workload (3 replacements instead of 4)")
    ap.add_argument("--reliable-interval", type=int, default=0,
                    help="FP32-refined CG: iterations between FP64 residual replacements (0 = library default)")
    ap.add_argument("--nu", type=int, default=2, help="pre- and post-smoothing sweeps")
    ap.add_argument("--sa-omega", type=float, default=2.0 / 3.0, help="prolongator smoothing weight")
    ap.add_argument("--min-coarse", type=int, default=None, help="stop coarsening at this many rows")
    ap.add_argument("--unsmoothed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the SpMV roofline (and copy calibration): the run profiled for byte counters")
    ap.add_argument("--spmv-format", choices=["csr_int32", "auto", "both"], default="both",
                    help="--spmv-only: the device format(s) timed (csr_int32: row-block CSR, int32 columns; "
                         "auto: the format the solver uses)")
    ap.add_argument("--spmv-launches", type=int, default=200)
    ap.add_argument("--spmv-copies", type=int, default=4, help="device copies of the operator rotated in the SpMV timing")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the secondary workloads (2D Poisson, cavity, networks, adjoint)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also time device-to-device copies of 512 MiB (scales rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU sharding is described, not run",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import cfdrs_amd as ca
    from cfdrs_amd import problems as P
    from cfdrs_amd.sparse import bind_torch_stream

    dev = torch.device("cuda:0")
    nn = args.grid
    n = nn ** 3
    log(f"generating {nn}^3 problem")
    rp, ci, v, b_host = P.config2(nn)
    nnz = ci.size
    a = ca.CsrMatrix.from_parts(v, ci, rp, n, n)
    if args.spmv_only:
        bind_torch_stream()
        fmts = ["csr_int32", "auto"] if args.spmv_format == "both" else [args.spmv_format]
        res = {}
        for f in fmts:
            sp = spmv_roofline((v, ci, rp), n, nnz, b_host, args, f)
            res[f] = {"format": sp["format"], "spmv_fp64_us": round(sp[8]["us"], 3),
                      "spmv_fp64_gbs": round(sp[8]["gbs"], 1), "spmv_fp32_us": round(sp[4]["us"], 3),
                      "spmv_fp32_gbs": round(sp[4]["gbs"], 1), "algorithmic_bytes_fp64": sp[8]["bytes"],
                      "algorithmic_bytes_fp32": sp[4]["bytes"]}
        res["copy_calibration"] = copy_calibration(args.copy_calibration)
        print(json.dumps(res), flush=True)
        return

    cfg = ca.AMGConfig.smoothed_aggregation(args.theta, args.omega, smoothed=not args.unsmoothed)
    cfg.pre_smooth_iterations = cfg.post_smooth_iterations = args.nu
    cfg.smoothing_omega = args.sa_omega
    if args.min_coarse is not None:
        cfg.min_coarse_size = args.min_coarse
    log("AMG setup")
    amg = ca.AlgebraicMultigrid(a, cfg)
    st = amg.statistics
    log(f"levels {st.level_sizes} nnz {st.level_nnz} op-complexity {st.operator_complexity:.3f} "
        f"setup {st.setup_time:.2f}s")

    b = torch.from_numpy(b_host).to(dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    bind_torch_stream()

    def solve(precision):
        x.zero_()
        c = (ca.IterativeSolverConfig.new(1e-10).with_max_iterations(1000)
             .with_precision(precision, args.inner_tol).with_reliable_interval(args.reliable_interval)
             .with_reliable_delta(args.reliable_delta))
        return ca.cg_preconditioned(a, b, amg, x, c)

    results = {}
    for prec in (ca.Precision.FP64, ca.Precision.FP32_REFINED):
        for _ in range(args.warmup):
            solve(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = [solve(prec) for _ in range(args.steps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        r = torch.empty_like(x)
        rn = a_residual(a, b, x, r)
        rel = rn / float(torch.linalg.norm(b))
        rep = reps[-1]
        log(f"{prec.name}: {ms:.3f} ms/solve, {rep.iterations} iterations ({rep.outer_iterations} outer), "
            f"{rep.termination.name}, relative residual {rel:.3e}")
        if rep.termination != ca.Termination.Converged or rel > 1e-10:
            raise SystemExit(f"{prec.name} solve did not reach 1e-10: {rep}, residual {rel}")
        results[prec.name] = dict(ms=ms, iterations=rep.iterations, outer=rep.outer_iterations, rel_residual=rel,
                                  x=x.clone())

    err = float(torch.linalg.norm(results["FP32_REFINED"]["x"] - results["FP64"]["x"]) /
                torch.linalg.norm(results["FP64"]["x"]))
    auto_picks = solve(ca.Precision.AUTO).precision.name  # the default policy's choice here
    log(f"Precision.AUTO picks {auto_picks}")
    # solve-level bytes: the per-iteration model of the metric, B_spmv + B_prec
    # + 8 n s, times the iterations, over the measured time
    solve_bw = {}
    for name, s_ in (("FP64", 8), ("FP32_REFINED", 4)):
        per_it = pcg_iteration_bytes(amg, n, nnz, s_, cfg)
        its = results[name]["iterations"]
        solve_bw[name] = {"bytes_per_iteration": per_it, "iterations": its,
                          "achieved_gbs": round(per_it * its / (results[name]["ms"] * 1e-3) / 1e9, 1)}
        solve_bw[name]["frac_of_hbm_peak"] = round(solve_bw[name]["achieved_gbs"] / HBM_PEAK_GBS, 4)
    log(f"solve-level model bandwidth: {solve_bw}")

    # SpMV roofline: FP64 and FP32 y = A x on the fine operator, timed with
    # HIP events recorded on the stream the library launches on (torch's
    # current stream, bound above), as row-block CSR with int32 columns (the
    # layout of the metric's byte model) and in the solver's own format.  The
    # launches run back to back and rotate over --spmv-copies device copies
    # of the operator and of x/y, so each launch streams a matrix last
    # touched copies-1 launches earlier and the working set exceeds the
    # 256 MB Infinity Cache, as inside a solve, where the V-cycle streams the
    # other levels between two fine-level products.
    spmv = spmv_roofline((v, ci, rp), n, nnz, b_host, args, "csr_int32")
    spmv_dev = spmv_roofline((v, ci, rp), n, nnz, b_host, args, "auto")

    copy = copy_calibration(args.copy_calibration)

    # the same setup with the Galerkin products formed on the device
    cfg_dev = ca.AMGConfig(**{**cfg.__dict__, "galerkin_on_device": True})
    t0 = time.perf_counter()
    amg_dev = ca.AlgebraicMultigrid(a, cfg_dev)
    setup_dev_s = time.perf_counter() - t0
    same_levels = amg_dev.statistics.level_nnz == st.level_nnz
    amg_dev.close()
    log(f"AMG setup with device Galerkin products: {setup_dev_s:.2f}s (host products {st.setup_time:.2f}s), "
        f"same level sizes: {same_levels}")

    # AMG recompute with new fine values (same values here), on the device
    # against the host Galerkin products plus upload
    recompute = {"device_ms": round(_time_solves(lambda: amg.recompute_device(a), 3)[0], 3),
                 "host_ms": round(_time_solves(lambda: amg.recompute(a), 1, warmup=0)[0], 1)}
    log(f"AMG recompute {nn}^3: {recompute}")

    others = None if args.no_configs else secondary_configs(args)
    if others is not None and not args.no_cpu:
        others["cpu_reference"] = cpu_secondary(args)

    cpu = None
    oracle_err = None
    if not args.no_cpu:
        cpu, x_orc = cpu_baseline(rp, ci, v, b_host, n, args, cfg)
        # solution error of each GPU solve against the CPU FP64 oracle's
        # solution of the same system (both stopped at relative residual 1e-10)
        oracle_err = {name: float(np.linalg.norm(results[name]["x"].cpu().numpy() - x_orc) / np.linalg.norm(x_orc))
                      for name in ("FP64", "FP32_REFINED")}
        log(f"solution error against the CPU FP64 oracle: {oracle_err}")

    t64, tref = results["FP64"]["ms"], results["FP32_REFINED"]["ms"]
    line = {
        "metric": "time-to-1e-10 residual & SpMV HBM GB/s: FP32-refined vs FP64 GPU",
        "value": round(tref, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tref, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": "synthetic (seeded SplitMix64 coefficients and right-hand side)",
        "config": {"workload": f"3D 7-point variable-coefficient Poisson {nn}^3, SA-AMG-PCG to 1e-10",
                   "unknowns": n, "nnz": int(nnz), "amg_levels": st.level_sizes,
                   "operator_complexity": round(st.operator_complexity, 4), "amg_setup_s": round(st.setup_time, 3),
                   "amg_setup_device_galerkin_s": round(setup_dev_s, 3),
                   "parallelism": "single GPU"},
        "fp64_ms": round(t64, 4),
        "fp64_iterations": results["FP64"]["iterations"],
        "refined_iterations": results["FP32_REFINED"]["iterations"],
        "refined_outer_steps": results["FP32_REFINED"]["outer"],
        "speedup_refined_vs_fp64": round(t64 / tref, 4),
        "precision_auto_picks": auto_picks,
        "rel_residual_fp64": results["FP64"]["rel_residual"],
        "rel_residual_refined": results["FP32_REFINED"]["rel_residual"],
        "refined_vs_fp64_solution_rel_diff": err,
        "solution_rel_error_vs_cpu_fp64_oracle": oracle_err,
        "solve_level_bandwidth": solve_bw,
        "copy_calibration": copy,
        "amg_recompute": recompute,
        "roofline": {"bound": "hbm", "achieved": round(spmv[8]["gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv[8]["gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": SPMV_FP64_TRAFFIC_BYTES.get(spmv["format"]) if nn == 128 else None,
                     "kernel": "FP64 y = A x, " + spmv["kernel"],
                     "avg_us": round(spmv[8]["us"], 3),
                     "timing": f"HIP events over {args.spmv_launches} back-to-back launches rotating over "
                               f"{args.spmv_copies} device copies of the operator",
                     "algorithmic_bytes": spmv[8]["bytes"],
                     "same_bytes_copy_gbs": round(spmv[8]["copy_gbs"], 1),
                     "frac_of_same_bytes_copy": round(spmv[8]["gbs"] / spmv[8]["copy_gbs"], 4),
                     "fp32_us": round(spmv[4]["us"], 3), "fp32_gbs": round(spmv[4]["gbs"], 1),
                     "fp32_frac": round(spmv[4]["gbs"] / HBM_PEAK_GBS, 4)},
        "spmv_device_format": {"kernel": "FP64 y = A x, " + spmv_dev["kernel"], "avg_us": round(spmv_dev[8]["us"], 3),
                               "format_bytes": spmv_dev[8]["bytes"], "achieved_gbs": round(spmv_dev[8]["gbs"], 1),
                               "frac": round(spmv_dev[8]["gbs"] / HBM_PEAK_GBS, 4),
                               "traffic": SPMV_FP64_TRAFFIC_BYTES.get(spmv_dev["format"]) if nn == 128 else None,
                               "same_bytes_copy_gbs": round(spmv_dev[8]["copy_gbs"], 1),
                               "fp32_us": round(spmv_dev[4]["us"], 3), "fp32_gbs": round(spmv_dev[4]["gbs"], 1)},
        "cpu_baseline": cpu,
        "other_workloads": others,
    }
    print(json.dumps(line), flush=True)


def spmv_roofline(parts, n: int, nnz: int, b_host, args, fmt: str) -> dict:
    """Average duration and bandwidth of y = A x on the fine operator, FP64
    and FP32, with the device format restricted to fmt ("csr_int32": row-block
    CSR with int32 columns; "auto": the format the solver picks).  K launches
    are enqueued back to back between two HIP events on the library's stream,
    launch k on copy k % copies of the operator (and of x and y), preceded by
    one untimed pass over all copies."""
    import ctypes as C

    import torch

    import cfdrs_amd as ca
    from cfdrs_amd import _native as N

    v, ci, rp = parts
    dev = torch.device("cuda:0")
    mats = [ca.CsrMatrix.from_parts(v, ci, rp, n, n) for _ in range(args.spmv_copies)]
    for m in mats:
        storage = m.set_storage(fmt)
    if fmt == "csr_int32":
        storage = dict(storage, format="csr_int32")
    # FP64 takes the two-rows-per-lane kernel when n and the plane offsets are even
    dia_kernel = "k_dia_vec, two rows per lane" if args.grid % 2 == 0 else "k_dia"
    kernel = {"dia": f"{dia_kernel} (symmetric diagonal storage, {storage['units']} stored diagonals x {n} rows)",
              "sell64": f"k_sell (64-row slices, int16 offsets, {storage['stored_entries']} stored slots)",
              "csr_rowblock": f"k_csr_rb ({storage['units']} row blocks)",
              "csr_int32": f"k_csr_rb, row-block CSR with int32 row pointers and columns ({storage['units']} row "
                           f"blocks of at most 256 rows and 2048 nonzeros)"}[storage["format"]]
    if storage["format"] in ("csr_rowblock", "csr_int32"):
        # gather layout the upload picked for this operator, per value type
        g = mats[0].gather()
        kernel += f"; gathers: FP64 {g['fp64']}, FP32 {g['fp32']}"
    fn = N.lib().cfdla_spmv_device
    out = {"kernel": kernel, "format": storage["format"]}
    for dt, s, code in ((torch.float64, 8, N.DT_F64), (torch.float32, 4, N.DT_F32)):
        xs = [torch.from_numpy(b_host).to(dev).to(dt) for _ in mats]
        ys = [torch.empty_like(xs[0]) for _ in mats]
        calls = [(m.handle, code, C.c_void_p(xv.data_ptr()), C.c_void_p(yv.data_ptr()))
                 for m, xv, yv in zip(mats, xs, ys)]
        for c in calls:
            if fn(*c) != N.OK:
                raise RuntimeError(N.lib().cfdla_last_error().decode())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(args.spmv_launches):
            fn(*calls[k % len(calls)])
        e1.record()
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / args.spmv_launches
        nbytes = spmv_bytes(n, nnz, s, storage)
        out[s] = dict(us=us, bytes=nbytes, gbs=nbytes / (us * 1e-6) / 1e9)
        log(f"SpMV {storage['format']} fp{8 * s}: {us:.2f} us/launch ({out[s]['gbs']:.1f} GB/s of "
            f"{nbytes / 1e6:.2f} MB)")
        del xs, ys
        if args.spmv_only:  # the counter runs calibrate on the copy kernel: no other copies there
            continue
        # practical ceiling at this size: device copies moving the same total
        # bytes (half read, half written), rotated over as many buffer pairs
        srcs = [torch.ones(nbytes // 16, dtype=torch.float64, device=dev) for _ in mats]
        dsts = [torch.empty_like(t) for t in srcs]
        for q in range(len(srcs)):
            dsts[q].copy_(srcs[q])
        e0.record()
        for k in range(args.spmv_launches):
            dsts[k % len(srcs)].copy_(srcs[k % len(srcs)])
        e1.record()
        e1.synchronize()
        cus = e0.elapsed_time(e1) * 1e3 / args.spmv_launches
        out[s]["copy_us"] = cus
        out[s]["copy_gbs"] = nbytes / (cus * 1e-6) / 1e9
        log(f"  same-bytes device copy: {cus:.2f} us ({out[s]['copy_gbs']:.1f} GB/s); SpMV at "
            f"{us and cus / us:.3f} of it")
        del srcs, dsts
    for m in mats:
        m.close()
    return out


def copy_calibration(enabled: bool):
    """Device-to-device copy of 512 MiB, 3 launches timed with HIP events: the
    bandwidth a plain streaming kernel reaches, and the known-size launches
    that scale the profiler's byte counters."""
    if not enabled:
        return None
    import torch

    src = torch.ones(64 << 20, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / 3
    nbytes = 2 * src.numel() * 8
    log(f"copy calibration: {src.numel() * 8} bytes read and written per copy, {us:.1f} us each, "
        f"{nbytes / (us * 1e-6) / 1e9:.1f} GB/s")
    # read-only streaming: a sum over the two 512 MiB buffers alternately
    tot = torch.zeros((), dtype=torch.float64, device="cuda:0")
    tot += src.sum()
    e0.record()
    for k in range(4):
        tot += (src if k % 2 == 0 else dst).sum()
    e1.record()
    e1.synchronize()
    rus = e0.elapsed_time(e1) * 1e3 / 4
    rbytes = src.numel() * 8
    log(f"read calibration: sum over {rbytes} bytes, {rus:.1f} us each, {rbytes / (rus * 1e-6) / 1e9:.1f} GB/s")
    return {"bytes_per_launch": nbytes, "us": round(us, 2), "gbs": round(nbytes / (us * 1e-6) / 1e9, 1),
            "read_gbs": round(rbytes / (rus * 1e-6) / 1e9, 1)}


def _time_solves(fn, steps: int, warmup: int = 1):
    """Mean wall ms of fn() over steps calls (device synchronized), and its last result."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, out


def secondary_configs(args) -> dict:
    """Time-to-1e-10 on the four other workloads (inputs resident on the
    device, setup excluded): 2D pressure-correction Jacobi-PCG (128^2), the
    lid-driven cavity momentum (BiCGSTAB, GMRES(30)) and pressure (AMG-PCG)
    systems (512^2), 4096 batched Kirchhoff networks, and the adjoint
    gradient of the outlet flux on a 256^2 Gaussian-random-field
    permeability."""
    import numpy as np
    import torch

    import cfdrs_amd as ca
    from cfdrs_amd import adjoint as ADJ
    from cfdrs_amd import network as NW
    from cfdrs_amd import problems as P

    dev = torch.device("cuda:0")
    steps = 3
    out = {}

    def krylov(a, b_host, prec, kind, precision, restart=30):
        b = torch.from_numpy(b_host).to(dev)
        x = torch.zeros_like(b)
        cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(precision)

        def run():
            x.zero_()
            if kind == "cg":
                return ca.cg_preconditioned(a, b, prec, x, cfg)
            if kind == "bicgstab":
                return ca.bicgstab_preconditioned(a, b, prec, x, cfg)
            return ca.gmres_preconditioned(a, b, prec, x, cfg, restart)

        ms, rep = _time_solves(run, steps)
        return {"ms": round(ms, 3), "iterations": rep.iterations, "outer": rep.outer_iterations,
                "termination": rep.termination.name, "solves_per_s": round(1e3 / ms, 1)}

    def auto_pick(a, b_host, prec, kind, restart=30):
        # the precision Precision.AUTO chooses for this system (one solve)
        b = torch.from_numpy(b_host).to(dev)
        x = torch.zeros_like(b)
        cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(20000).with_precision(ca.Precision.AUTO)
        fn = {"cg": ca.cg_preconditioned, "bicgstab": ca.bicgstab_preconditioned}.get(kind)
        rep = fn(a, b, prec, x, cfg) if fn else ca.gmres_preconditioned(a, b, prec, x, cfg, restart)
        return rep.precision.name

    # 2D pressure correction, 128^2, Jacobi-PCG
    rp, ci, v, b1 = P.config1(128)
    a1 = ca.CsrMatrix.from_parts(v, ci, rp, 128 * 128, 128 * 128)
    j1 = ca.JacobiPreconditioner(a1)
    out["pressure_correction_128x128_jacobi_pcg"] = {
        prec.name: krylov(a1, b1, j1, "cg", prec) for prec in (ca.Precision.FP64, ca.Precision.FP32_REFINED)}
    out["pressure_correction_128x128_jacobi_pcg"]["AUTO_picks"] = auto_pick(a1, b1, j1, "cg")
    log(f"2D pressure correction: {out['pressure_correction_128x128_jacobi_pcg']}")

    # lid-driven cavity 512^2: momentum (nonsymmetric) and pinned pressure correction
    (mrp, mci, mv, bm), (prp, pci, pv, bp) = P.config3(512)
    am = ca.CsrMatrix.from_parts(mv, mci, mrp, 512 * 512, 512 * 512)
    jm = ca.JacobiPreconditioner(am)
    npress = bp.size
    ap_ = ca.CsrMatrix.from_parts(pv, pci, prp, npress, npress)
    amg_p = ca.AlgebraicMultigrid(ap_, ca.AMGConfig.smoothed_aggregation())
    sm = ca.SuccessiveOverRelaxation(am, 1.0)  # the reference's momentum preconditioner, red-black order
    cav = {}
    for prec in (ca.Precision.FP64, ca.Precision.FP32_REFINED):
        cav[f"momentum_bicgstab_jacobi_{prec.name}"] = krylov(am, bm, jm, "bicgstab", prec)
        cav[f"momentum_gmres30_jacobi_{prec.name}"] = krylov(am, bm, jm, "gmres", prec)
        cav[f"momentum_gmres30_sor_{prec.name}"] = krylov(am, bm, sm, "gmres", prec)
        cav[f"pressure_amg_pcg_{prec.name}"] = krylov(ap_, bp, amg_p, "cg", prec)
    cav["AUTO_picks"] = {"momentum_bicgstab_jacobi": auto_pick(am, bm, jm, "bicgstab"),
                         "momentum_gmres30_jacobi": auto_pick(am, bm, jm, "gmres"),
                         "momentum_gmres30_sor": auto_pick(am, bm, sm, "gmres"),
                         "pressure_amg_pcg": auto_pick(ap_, bp, amg_p, "cg")}
    out["lid_driven_cavity_512x512_re1000"] = cav
    log(f"cavity: {cav}")

    # 4096 Kirchhoff networks of 50..2000 nodes, one persistent launch
    nets = NW.synth_networks()
    batch = NW.channel_batch(nets)
    xb = torch.zeros(batch.n_rows, dtype=torch.float64, device=dev)
    netres = {"systems": nets.count, "rows": int(batch.n_rows)}
    # per-system rows and nonzeros, for the bytes each CG iteration streams
    net_rp = batch.export()[0]
    net_n = np.diff(nets.node_offset).astype(np.float64)
    net_nnz = (net_rp[nets.node_offset[1:]] - net_rp[nets.node_offset[:-1]]).astype(np.float64)
    for prec in (ca.Precision.FP64, ca.Precision.FP32_REFINED):
        cfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(50000).with_precision(prec)
        # the row-bin schedule the autotuner picks for this size histogram
        # (timed once per candidate, cached as JSON)
        tuned = batch.autotune(cfg)

        def run():
            xb.zero_()
            return batch.solve(xb, cfg)[1]

        ms, reps = _time_solves(run, steps)
        # Each team CG iteration re-reads its system's matrix (values, int32
        # columns and row pointers) and the Jacobi and equilibration factors:
        # nnz (s + 4) + 4 (n + 1) + n (s + 8) bytes, s = 8 (FP64) / 4 (FP32
        # inner iterations); the FP32 form's FP64 residual replacements are
        # not counted.  Most re-reads hit the L2 of the team's XCD, so this is
        # an L2-served stream rate, not an HBM bandwidth.
        sv = 8 if prec == ca.Precision.FP64 else 4
        its = np.array([r.iterations for r in reps], np.float64)
        stream = float(np.sum(its * (net_nnz * (sv + 4) + 4 * (net_n + 1) + net_n * (sv + 8))))
        netres[prec.name] = {"ms": round(ms, 3), "systems_per_s": round(nets.count * 1e3 / ms, 1),
                             "max_iterations": int(max(r.iterations for r in reps)),
                             "mean_iterations": round(float(its.mean()), 1),
                             "all_converged": all(r.termination == ca.Termination.Converged for r in reps),
                             "matrix_stream_gb": round(stream / 1e9, 2),
                             "l2_served_matrix_stream_gbs": round(stream / (ms * 1e-3) / 1e9, 1),
                             "schedule": tuned["best"]}
    # design-loop refresh: widths -> conductances -> assembled values, right-
    # hand side and preconditioner factors, on the device against the host
    w_dev = torch.from_numpy(nets.width).to(dev)
    l_dev = torch.from_numpy(nets.length).to(dev)

    def refresh_device():
        g, _ = NW.rect_conductance_device(w_dev, NW.CHANNEL_HEIGHT, l_dev)
        batch.update_conductance_device(g)

    def refresh_host():
        g, _ = NW.rect_conductance(nets.width, NW.CHANNEL_HEIGHT, nets.length)
        batch.update_conductance(g)

    ms_dev, _ = _time_solves(refresh_device, steps)
    ms_host, _ = _time_solves(refresh_host, 1, warmup=0)
    netres["refresh_device_ms"] = round(ms_dev, 3)
    netres["refresh_host_ms"] = round(ms_host, 1)
    netres["edges"] = int(nets.edge_i.size)
    out["networks_4096"] = netres
    log(f"networks: {netres}")

    # SIMPLE pressure-correction step on a 512^2 channel grid (walls south
    # and north, inflow west, pressure outlet east; solvers/simple/pressure.rs):
    # matrix and right-hand side refreshed from the fields, AMG hierarchy
    # recomputed, BiCGSTAB-AMG to 1e-10, corrections -- all on the device.
    # The host path of the same refresh (CPU assembly, values over PCIe, host
    # Galerkin recompute) is timed beside it; its solve would be the same.
    from cfdrs_amd import simple as S

    nx = ny = 512
    ns = nx * ny
    hs = 1.0 / (nx - 1)
    fields = [P.uniform(301, ns), P.uniform(302, ns), P.uniform(303, ns), 1e-2 * P.uniform(304, ns, 0.5, 1.5),
              1e-2 * P.uniform(305, ns, 0.5, 1.5), np.ones(ns)]
    side = (S.SIDE_PLAIN, S.SIDE_PRESSURE, S.SIDE_PLAIN, S.SIDE_PLAIN)
    sysm = S.SimplePressureCorrection(nx, ny, hs, hs, side, fields=fields, positive_diagonal=True)
    amg_s = ca.AlgebraicMultigrid(sysm.matrix, ca.AMGConfig.smoothed_aggregation())
    fd = [torch.from_numpy(f).to(dev) for f in fields]
    pc, uc, vc = fd[2].clone(), fd[0].clone(), fd[1].clone()  # corrected copies keep the system stationary
    rhs_s = torch.empty(ns, dtype=torch.float64, device=dev)
    pp = torch.zeros(ns, dtype=torch.float64, device=dev)
    scfg = ca.IterativeSolverConfig.new(1e-10).with_max_iterations(2000)

    def s_refresh():
        sysm.assemble_device(*fd, rhs_s, want_max=False)

    def s_recompute():
        amg_s.recompute_device(sysm.matrix)

    def s_solve():
        pp.zero_()
        return ca.bicgstab_preconditioned(sysm.matrix, rhs_s, amg_s, pp, scfg)

    def s_correct():
        sysm.apply_corrections_device(pp, fd[3], fd[4], pc, uc, vc, 0.3)

    def s_step():
        s_refresh()
        s_recompute()
        rep = s_solve()
        s_correct()
        return rep

    ms_step, rep_s = _time_solves(s_step, steps)
    simple = {"grid": f"{nx}x{ny}", "unknowns": ns, "form": "positive diagonal (interior rows negated)",
              "amg_levels": amg_s.statistics.level_sizes,
              "step_ms": round(ms_step, 3), "iterations": rep_s.iterations,
              "termination": rep_s.termination.name}
    for name, fn in (("refresh_ms", s_refresh), ("amg_recompute_device_ms", s_recompute), ("solve_ms", s_solve),
                     ("corrections_ms", s_correct)):
        simple[name] = round(_time_solves(fn, steps)[0], 3)
    # the reference's sign (a_P = -sum a_nb beside +1 boundary rows): its
    # own solver, unpreconditioned BiCGSTAB (pressure.rs:452-485), and the
    # same AMG recipe
    ref_sys = S.SimplePressureCorrection(nx, ny, hs, hs, side, fields=fields)
    rhs_r = torch.empty(ns, dtype=torch.float64, device=dev)
    ref_sys.assemble_device(*fd, rhs_r, want_max=False)
    amg_r = ca.AlgebraicMultigrid(ref_sys.matrix, ca.AMGConfig.smoothed_aggregation())

    def r_plain():
        pp.zero_()
        return ca.bicgstab(ref_sys.matrix, rhs_r, pp, scfg)

    def r_amg():
        pp.zero_()
        return ca.bicgstab_preconditioned(ref_sys.matrix, rhs_r, amg_r, pp, scfg)

    for name, fn in (("reference_sign_bicgstab_unpreconditioned", r_plain), ("reference_sign_bicgstab_amg", r_amg)):
        ms_r, rep_r = _time_solves(fn, 1)
        simple[name] = {"ms": round(ms_r, 3), "iterations": rep_r.iterations, "termination": rep_r.termination.name}
    a_h = ca.CsrMatrix.from_parts(sysm.matrix.values, sysm.matrix.col_indices, sysm.matrix.row_ptr, ns, ns)
    amg_h = ca.AlgebraicMultigrid(a_h, ca.AMGConfig.smoothed_aggregation())
    amg_h.apply_to(np.zeros(ns), np.empty(ns))  # hierarchy on the device, as in the device path

    def h_assemble():
        return P.simple_pressure(nx, ny, hs, hs, side, *fields, positive_diagonal=True)[2]

    vals = h_assemble()
    simple["host_assembly_ms"] = round(_time_solves(h_assemble, 1, warmup=0)[0], 3)
    simple["host_upload_ms"] = round(_time_solves(lambda: a_h.update_values(vals), 1, warmup=0)[0], 3)
    simple["host_amg_recompute_ms"] = round(_time_solves(lambda: amg_h.recompute(a_h), 1, warmup=0)[0], 3)
    out["simple_pressure_step_512x512_channel"] = simple
    log(f"SIMPLE pressure step: {simple}")

    # adjoint gradient of the outlet flux, 256^2 GRF permeability: the first
    # call builds the AMG hierarchy on the host; later calls assemble on the
    # device, recompute the hierarchy's Galerkin products on the device, and
    # run the forward and adjoint solves and the gradient kernel
    prob = ADJ.FluxProblem(256, 256)
    g = torch.from_numpy(P.grf(256, 256)).to(dev)
    first_ms, _ = _time_solves(lambda: ADJ.flux_and_gradient(prob, g), 1, warmup=0)
    fwd_ms, _ = _time_solves(lambda: prob.forward(g), steps)
    ms, _ = _time_solves(lambda: ADJ.flux_and_gradient(prob, g), steps)
    out["adjoint_flux_gradient_256x256"] = {"ms": round(ms, 3), "forward_ms": round(fwd_ms, 3),
                                            "first_call_ms_including_amg_setup": round(first_ms, 3)}
    log(f"adjoint: {out['adjoint_flux_gradient_256x256']}")
    return out


def cpu_secondary(args) -> dict:
    """The CPU FP64 oracle (the reference's algorithms restated) on the four
    other workloads, timed on the host's cores (up to 16 threads) and on one
    thread, setup and assembly excluded: 2D pressure-correction Jacobi-PCG
    (128^2), cavity momentum BiCGSTAB and GMRES(30) with Jacobi and the pinned
    pressure AMG-PCG (512^2), the reference's per-network cascade
    (linear_system.rs:114-267: row equilibration, dense LU up to 256 rows,
    else Jacobi-CG and its LU fallback) over a seeded sample of the 4096
    networks, and the forward + adjoint solves and gradient on the 256^2
    permeability field."""
    import concurrent.futures as cf

    import numpy as np

    from cfdrs_amd import network as NW
    from cfdrs_amd import problems as P
    from oracle import oracle as O

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    out = {"cores": threads, "kind": "port", "host": host_cpu()}
    log(f"CPU oracle on the secondary workloads ({threads} threads, then 1)")

    def timed(fn):
        res = {}
        for nt in (threads, 1):
            O.set_threads(nt)
            t0 = time.perf_counter()
            extra = fn()
            res["ms" if nt == threads else "single_thread_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
            res.update(extra or {})
        O.set_threads(1)
        return res

    rp, ci, v, b1 = P.config1(128)
    a1 = O.Csr.from_arrays(rp, ci, v, 128 * 128)
    out["pressure_correction_128x128_jacobi_pcg"] = timed(
        lambda: {"iterations": O.solve(a1, b1, kind="cg", prec="jacobi", tol_rel=1e-10, max_it=20000)[1].iterations})

    (mrp, mci, mv, bm), (prp, pci, pv, bp) = P.config3(512)
    am = O.Csr.from_arrays(mrp, mci, mv, 512 * 512)
    ap_ = O.Csr.from_arrays(prp, pci, pv, bp.size)
    cav = {}
    cav["momentum_bicgstab_jacobi"] = timed(
        lambda: {"iterations": O.solve(am, bm, kind="bicgstab", prec="jacobi", tol_rel=1e-10,
                                       max_it=20000)[1].iterations})
    cav["momentum_gmres30_jacobi"] = timed(
        lambda: {"iterations": O.solve(am, bm, kind="gmres", prec="jacobi", tol_rel=1e-10, max_it=20000,
                                       restart=30)[1].iterations})
    amg_p = O.Amg(ap_, coarsening="three_phase", theta=0.25, interp="smoothed", smoother="jacobi", omega=0.8,
                  coarse_solve="direct", min_coarse=1000)
    cav["pressure_amg_pcg"] = timed(
        lambda: {"iterations": O.solve(ap_, bp, kind="cg", prec="amg", amg=amg_p, tol_rel=1e-10,
                                       max_it=1000)[1].iterations})
    out["lid_driven_cavity_512x512_re1000"] = cav

    nets = NW.synth_networks()
    batch = NW.channel_batch(nets)
    nrp, nci, nv, nb = batch.export()
    no = nets.node_offset
    sample = np.sort(np.random.default_rng(4999).choice(nets.count, size=512, replace=False))
    systems = []
    for s_ in sample:
        r0, r1 = int(no[s_]), int(no[s_ + 1])
        k0, k1 = int(nrp[r0]), int(nrp[r1])
        systems.append((O.Csr.from_arrays(nrp[r0:r1 + 1] - k0, nci[k0:k1] - r0, nv[k0:k1], r1 - r0),
                        np.ascontiguousarray(nb[r0:r1])))
    O.set_threads(1)

    def cascade(sys_):
        return O.network_cascade(sys_[0], sys_[1], tol=1e-10, max_it=1000)[1]

    net = {"sample": f"{sample.size} of the {nets.count} networks (seeded uniform draw, "
                     f"{int(sum(no[s_ + 1] - no[s_] for s_ in sample))} nodes), equilibrated tolerance 1e-10, "
                     f"1000 CG iterations before the LU fallback (the reference's default cap)"}
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as pool:  # one network per task; ctypes releases the GIL
        paths = list(pool.map(cascade, systems))
    ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for sys_ in systems:
        cascade(sys_)
    ms1 = (time.perf_counter() - t0) * 1e3
    net.update({"ms": round(ms, 2), "single_thread_ms": round(ms1, 2),
                "systems_per_s": round(sample.size * 1e3 / ms, 1),
                "single_thread_systems_per_s": round(sample.size * 1e3 / ms1, 1),
                "paths": {name: int(sum(1 for p_ in paths if p_ == code))
                          for code, name in ((0, "dense"), (1, "krylov"), (2, "dense_fallback"))}})
    out["networks_4096"] = net

    from cfdrs_amd import adjoint as ADJ

    nx = ny = 256
    g = P.grf(nx, ny)
    k = np.exp(g)
    a5, rhs5 = O.assemble_varcoef(nx, ny, 1, float(nx * nx), k, ADJ.DIR_MASK, np.array(ADJ.DIR_VAL))
    amg5 = O.Amg(a5, coarsening="three_phase", theta=0.25, interp="smoothed", smoother="jacobi", omega=0.8,
                 coarse_solve="direct", min_coarse=1000)
    east = np.arange(ny) * nx + nx - 1

    def adjoint():
        p_, _ = O.solve(a5, rhs5, kind="cg", prec="amg", amg=amg5, tol_rel=1e-10, max_it=1000)
        dJdp = np.zeros(nx * ny)
        dJdp[east] = k[east]
        lam, _ = O.solve(a5, dJdp, kind="cg", prec="amg", amg=amg5, tol_rel=1e-10, max_it=1000)
        dJdk = np.zeros(nx * ny)
        dJdk[east] = p_[east]
        O.coef_grad_2d(nx, ny, float(nx * nx), ADJ.DIR_MASK, np.array(ADJ.DIR_VAL[:4]), k, p_, lam, dJdk)

    out["adjoint_flux_gradient_256x256"] = timed(adjoint)
    out["adjoint_flux_gradient_256x256"]["sample"] = "forward and adjoint AMG-PCG solves to 1e-10 and the gradient"
    log(f"CPU oracle, secondary workloads: {out}")
    return out


def a_residual(a, b, x, r) -> float:
    import ctypes as C

    from cfdrs_amd import _native as N

    out = C.c_double()
    N.call("cfdla_residual_device", a.handle, b.data_ptr(), x.data_ptr(), r.data_ptr(), C.byref(out))
    return out.value


def cpu_baseline(rp, ci, v, b_host, n, args, cfg):
    """The CPU FP64 oracle solving the same system with the same AMG recipe:
    setup excluded, solve from x = 0 to 1e-10, on all host threads (at most
    16, the host share of one GPU) and on one thread."""
    import numpy as np

    from oracle import oracle as O

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    log(f"CPU oracle baseline ({threads} threads, then 1)")
    orc = O.Csr.from_arrays(rp, ci, v, n)
    t0 = time.perf_counter()
    amg = O.Amg(orc, coarsening="three_phase", theta=cfg.strength_threshold,
                interp="direct" if args.unsmoothed else "smoothed", smoother="jacobi",
                omega=cfg.relaxation_factor, coarse_solve="direct", sa_omega=cfg.smoothing_omega,
                min_coarse=cfg.min_coarse_size, pre=cfg.pre_smooth_iterations, post=cfg.post_smooth_iterations)
    setup = time.perf_counter() - t0
    timings = {}
    for nt in (threads, 1):
        O.set_threads(nt)
        t0 = time.perf_counter()
        x, rep = O.solve(orc, b_host, kind="cg", prec="amg", amg=amg, tol_rel=1e-10, max_it=1000)
        timings[nt] = time.perf_counter() - t0
        rel = float(np.linalg.norm(b_host - orc.spmv(x)) / np.linalg.norm(b_host))
        log(f"CPU ({nt} threads): setup {setup:.2f}s, solve {timings[nt]:.2f}s, {rep.iterations} iterations, "
            f"residual {rel:.2e}")
    O.set_threads(1)
    return {"value": round(timings[threads] * 1e3, 2), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"full FP64 AMG-PCG solve to 1e-10 of the same {n}-unknown system ({rep.iterations} "
                      f"iterations, relative residual {rel:.2e}); AMG setup {setup:.1f} s (one thread) excluded",
            "single_thread_ms": round(timings[1] * 1e3, 2), "iterations": rep.iterations,
            "host": host_cpu()}, x


def host_cpu() -> dict:
    """CPU model, physical cores and logical CPUs of this host (/proc/cpuinfo),
    and the CPUs this process may run on."""
    model, phys, logical = None, set(), 0
    try:
        pid = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, val = line.partition(":")
                k, val = k.strip(), val.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model is None:
                    model = val
                elif k == "physical id":
                    pid = val
                elif k == "core id":
                    core = val
                elif not k and pid is not None and core is not None:
                    phys.add((pid, core))
                    pid = core = None
        if pid is not None and core is not None:
            phys.add((pid, core))
    except OSError:
        pass
    return {"model": model, "physical_cores": len(phys) or None, "logical_cpus": logical or None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def pcg_iteration_bytes(amg, n: int, nnz: int, s: int, cfg) -> int:
    """Algorithmic bytes of one AMG-PCG iteration: B_spmv + B_prec + 8 n s,
    B_prec summing over the levels (nu1 + nu2) Jacobi sweeps (B_spmv,l +
    3 n_l s), the residual (B_spmv,l + n_l s), R_l and P_l products and
    2 n_l s, plus the coarsest dense solve (s_c n_c^2 + 2 n_c s, FP64 inverse)."""
    st = amg.statistics
    sizes, nnzs = st.level_sizes, st.level_nnz

    def spmv(rows, cols, z):
        return z * (s + 4) + rows * (s + 4) + cols * s + 4

    nu = cfg.pre_smooth_iterations + cfg.post_smooth_iterations
    prec = 0
    for l in range(len(sizes) - 1):
        nl = sizes[l]
        bl = spmv(nl, nl, nnzs[l])
        r_nnz = amg.level(l, "R")[1].size
        p_nnz = amg.level(l, "P")[1].size
        prec += nu * (bl + 3 * nl * s) + (bl + nl * s) + spmv(sizes[l + 1], nl, r_nnz) \
            + spmv(nl, sizes[l + 1], p_nnz) + 2 * nl * s
    nc = sizes[-1]
    prec += 8 * nc * nc + 2 * nc * s
    return int(spmv(n, n, nnz) + prec + 8 * n * s)


if __name__ == "__main__":
    main()
