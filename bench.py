#!/usr/bin/env python3
"""Global-placement benchmark of the MI355X engine.

Workload: a seeded synthetic 1M-cell design in the std-cell recipe (widths
lognormal median 3 sites, unit height, 1K boundary IO pins, 1.05M nets with
degree 2:55% 3:20% 4-10:20% 11-200 power-law, index-local pins, utilisation
0.7, target density 1.0; the reference's bin rule gives 1024x1024 bins).
A step is one accepted Nesterov iteration of global placement (the plain
backtracking method of nesterov.hpp:310-369 by default) with every state
array resident in HBM.  `value` = K timed steps / wall time.  The same run
also reports the wall time from the first gradient evaluation to overflow
<= 0.10 for plain backtracking and for the learned controller
(eplace_amd/controller_params.npy), and the CPU baseline: the oracle's restatement of the reference path
(reference float arithmetic, one thread) on a bounded sample of the same
design.

    python bench.py [--steps K] [--warmup W] [--method bktrk|vanilla]
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix


def device_sync():
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def up64(v):
    return -(-v // 64) * 64


def poisson_flops(info):
    """Multiply-add flops of the four parity-folded transform GEMM launches."""
    X, Y = up64(info["bins_x"]), up64(info["bins_y"])
    hx, hy = up64(info["bins_x"] // 2), up64(info["bins_y"] // 2)
    return (2.0 * X * hy * hy * 2 + 2.0 * hx * (2 * hy) * hx * 2 + 2.0 * (2 * hx) * hy * (2 * hy) * 2
            + 2.0 * hx * Y * (2 * hx) * 2)


# Algorithmic HBM bytes of one evaluation's phase (the bytes the operation
# must move once, independent of how it is implemented).
PHASE_BYTES = {
    # per object: x, y, w, h, kind, big flag (18 B); per bin: clear and accumulate two int64 maps
    "scatter": lambda O, N, P, nb: 18.0 * O + 32.0 * nb,
    # per bin: two int64 maps, terminal and base maps in, rho out
    "finalize": lambda O, N, P, nb: 28.0 * nb,
    # per pin: object index (4), offset (8), gathered centre (8), gradient out (8)
    "wirelength": lambda O, N, P, nb: 28.0 * P,
    # per pin: gradient slot index (4) and gathered gradient (8); per node: pin range (4), sum out (8)
    "node_sums": lambda O, N, P, nb: 12.0 * P + 12.0 * N,
    # per object: sorted index (4), footprint (16), gradient out (8); per bin: Ex, Ey once
    "field_gather": lambda O, N, P, nb: 28.0 * O + 8.0 * nb,
    # per object: kind, stage flag, pin count, area, density and wirelength gradients, previous
    # point and gradient, current point in (50 B), total gradient out (8 B)
    "combine": lambda O, N, P, nb: 58.0 * O,
}

PHASE_KERNELS = {
    "scatter": "tile_hist_kernel + tile_keyscan_kernel + tile_place_kernel + tile_scatter_kernel + scatter_big_kernel",
    "finalize": "finalize_density_kernel + finish_density_kernel",
    "wirelength": "wl_group_kernel<0,0,1> + wl_large_kernel<0,0,1>",
    "node_sums": "reduce_wl_kernel + node_wl_kernel",
    "field_gather": "gather_tile_kernel + gather_big_kernel",
    "combine": "gather_combine_kernel (after the second stream joins)",
}


def copy_peak_gbs(mib=1024, reps=10):
    """Bandwidth of a plain device-to-device copy (read + write bytes per
    second), timed with HIP events on the null stream: the practical HBM
    ceiling next to the datasheet peak."""
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    ms = ctypes.c_float()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    try:
        hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
        hip.hipEventCreate(ctypes.byref(a))
        hip.hipEventCreate(ctypes.byref(b))
        kind = 3  # hipMemcpyDeviceToDevice
        hip.hipMemcpyAsync(dst, src, ctypes.c_size_t(nbytes), kind, None)
        hip.hipEventRecord(a, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(dst, src, ctypes.c_size_t(nbytes), kind, None)
        hip.hipEventRecord(b, None)
        hip.hipEventSynchronize(b)
        hip.hipEventElapsedTime(ctypes.byref(ms), a, b)
        hip.hipEventDestroy(a)
        hip.hipEventDestroy(b)
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return 2.0 * nbytes / (ms.value * 1e-3 / reps) / 1e9


def phase_traffic(kernel_key, info):
    """HBM bytes of one evaluation's phase (all its launches) from the
    committed rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by a
    plain copy of known size), when it was measured on this workload;
    otherwise None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_per_launch.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    w = t.get("workload", {})
    if w.get("num_objects") != info["num_objects"] or w.get("bins") != [info["bins_x"], info["bins_y"]]:
        return None
    v = t.get("phases", {}).get(kernel_key)
    return None if v is None else float(v["bytes"])


def per_launch(v, launches):
    return None if v is None else v / launches


def roofline(name, seconds, info, poisson_mode=0):
    """Achieved rate of one evaluation phase against its roofline: HBM bytes
    for the streaming phases, MFMA flops for the Poisson transforms."""
    O, N, P = info["num_objects"], info["num_nodes"], info["num_pins"]
    nb = info["bins_x"] * info["bins_y"]
    if name == "poisson":
        flops = poisson_flops(info)
        if poisson_mode == 1:  # three bf16 products per FP32 product
            flops, peak, kern = 3 * flops, BF16_MFMA_PEAK_TFLOPS, "pgemm_kernel<1> x4 (split-BF16)"
        else:
            peak, kern = FP32_MFMA_PEAK_TFLOPS, "pgemm_kernel<0> x4 (FP32)"
        ach = flops / seconds / 1e12
        return {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                "traffic": per_launch(phase_traffic(name, info), 4), "kernel": kern, "launches": 4,
                "work_per_launch": flops / 4, "seconds_per_launch": seconds / 4}
    byts = PHASE_BYTES[name](O, N, P, nb)
    ach = byts / seconds / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": phase_traffic(name, info), "kernel": PHASE_KERNELS[name], "launches": 1,
            "work_per_launch": byts, "seconds_per_launch": seconds}


def cpu_baseline(design, method, iters, init_positions):
    from oracle.oracle import Oracle
    o = Oracle(design, 1.0, mode="ref")
    o.set("positions", init_positions.reshape(-1).astype(np.float64))
    o.run("total_gradient_update", 0, -1)
    o.run("penalty_init")
    o.run("opt_init")
    t0 = time.perf_counter()
    for _ in range(iters):
        o.run("opt_step", 1 if method == "bktrk" else 0, -1)
    dt = time.perf_counter() - t0
    return {"value": iters / dt, "unit": "iters/s", "cores": 1, "kind": "port",
            "sample": f"{iters} {method} Nesterov iterations of the same {design.num_nodes}-cell design after "
                      f"initialisation, oracle restatement of the reference path in reference float arithmetic, "
                      f"1 thread on {cpu_model()}",
            "seconds": dt}


def qp_reference_baseline(design, outer_iterations):
    """The reference's own QPPlacer (oracle/_ref/ref_harness, built from the
    reference's sources with its vendored Eigen, 8 OpenMP threads as
    qplace.cpp:50 asks) on the same design for `outer_iterations` outer
    iterations from the core centre; its per-iteration wall time is the
    baseline.  None when the harness is not built."""
    import subprocess
    import tempfile
    from eplace_amd.design import write_bookshelf
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not os.path.exists(exe):
        return None
    tmp = tempfile.mkdtemp(prefix="qpref_")
    try:
        aux = write_bookshelf(design, os.path.join(tmp, "in"))
        out = os.path.join(tmp, "out")
        env = dict(os.environ, OMP_NUM_THREADS="8")
        r = subprocess.run([exe, "qp", "-noQP", "-timeOnly", "-aux", aux, "-IPiteCount",
                            str(outer_iterations - 1), "-out", out], capture_output=True, text=True, cwd=tmp,
                           env=env, timeout=600)
        if r.returncode != 0:
            return None
        per = [float(l.split("CPUtime")[1]) for l in r.stdout.splitlines() if "at iteration number" in l]
        hp = [float(l.split("HPWL")[1].split(",")[0]) for l in r.stdout.splitlines() if "at iteration number" in l]
        if not per:
            return None
        return {"value": float(np.mean(per)), "unit": "s per outer iteration", "cores": 8, "kind": "reference",
                "sample": f"{len(per)} outer iteration(s) of the reference QPPlacer::quadraticPlacement "
                          f"(qplace.cpp, Eigen BiCGSTAB, setNbThreads(8)) on the same design from the core "
                          f"centre, wall time per iteration as the reference prints it, on {cpu_model()}",
                "hpwl_log": hp}
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--method", default="bktrk", choices=["bktrk", "vanilla"])
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-controller convergence run")
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--no-qp", action="store_true", help="skip the quadratic-placement run")
    ap.add_argument("--no-qp-cpu", action="store_true", help="skip the reference QP CPU baseline")
    ap.add_argument("--poisson-mode", type=int, default=1, choices=[0, 1],
                    help="transform arithmetic: 0 FP32 MFMA, 1 split-BF16 MFMA")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (scales HBM counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs global placement on one GPU; multi-GPU runs are not built "
              "(see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))
        return

    from eplace_amd import EPlacer2D, EplaceNesterovOpt, synthetic

    cells = args.cells
    nets = int(round(cells * 1.05))
    t = time.perf_counter()
    d = synthetic(cells, nets, seed=args.seed, num_io=1000, util=0.7, name=f"synth_{cells}")
    t_gen = time.perf_counter() - t
    print(f"[bench] design {cells} cells, {nets} nets, {d.num_pins} pins (generated in {t_gen:.1f}s)",
          file=sys.stderr, flush=True)
    t = time.perf_counter()
    gp = EPlacer2D(d, 1.0, poisson_mode=args.poisson_mode)
    info = gp.info()
    t_create = time.perf_counter() - t
    print(f"[bench] engine ready in {t_create:.1f}s: {info['num_objects']} objects "
          f"({info['num_fillers']} fillers), {info['bins_x']}x{info['bins_y']} bins", file=sys.stderr, flush=True)
    init_pos = gp.read("positions").copy()

    def restart():
        gp.write("positions", init_pos)
        gp.initialization()

    # warmup: W iterations from the initial placement (discarded)
    gp.initialization()
    EplaceNesterovOpt(gp, method=args.method, max_iterations=max(args.warmup - 1, 0)).opt()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # timed: exactly K iterations from the initial placement (evaluations
    # replayed from their captured graphs)
    restart()
    device_sync()
    t0 = time.perf_counter()
    r = EplaceNesterovOpt(gp, method=args.method, max_iterations=args.steps - 1, log=False).opt()
    device_sync()
    dt = time.perf_counter() - t0
    steps = r["iterations"]
    print(f"[bench] timed {steps} iterations in {dt:.3f}s, {r['evaluations']} evaluations, "
          f"{r['backtracks']} backtracks", file=sys.stderr, flush=True)
    # phase times: the same K iterations from the same start once more, every
    # evaluation launched kernel by kernel with HIP events between its phases
    # on the streams they run on (the events cost launches, so this pass is
    # not the timed one; its iterates are identical)
    restart()
    device_sync()
    t0 = time.perf_counter()
    r = EplaceNesterovOpt(gp, method=args.method, max_iterations=args.steps - 1, log=False).opt(profile=True)
    device_sync()
    dt_events = time.perf_counter() - t0
    ph = r["phase_seconds"]
    # dominant phase: the largest measured time per evaluation among the
    # kernel phases (the combine also holds the wait for the second stream)
    kernel_phases = [k for k in ph if k != "combine"]
    dominant = max(kernel_phases, key=ph.get)
    roof = roofline(dominant, ph[dominant], info, args.poisson_mode)
    roof_all = {k: roofline(k, ph[k], info, args.poisson_mode) for k in ph}

    conv = None
    if not args.no_converge:
        restart_t0 = time.perf_counter()
        gp.write("positions", init_pos)
        device_sync()
        t0 = time.perf_counter()
        gp.initialization()
        rc = EplaceNesterovOpt(gp, method=args.method, target_overflow=0.10, log=False).opt()
        device_sync()
        tc = time.perf_counter() - t0
        conv = {"seconds_to_overflow_0.10": tc, "iterations": rc["iterations"], "evaluations": rc["evaluations"],
                "backtracks": rc["backtracks"], "final_overflow": rc["final_overflow"],
                "final_hpwl": rc["final_hpwl"], "iters_per_s": rc["iterations"] / tc}
        print(f"[bench] converged: {rc['iterations']} iterations in {tc:.3f}s, overflow {rc['final_overflow']:.4f}, "
              f"HPWL {rc['final_hpwl']:.6g}", file=sys.stderr, flush=True)
        del restart_t0
        from eplace_amd.legalize import check_legal, legalize_abacus
        t0 = time.perf_counter()
        lxy, lst = legalize_abacus(d, gp.read("positions")[: d.num_nodes], mode=0)
        chk = check_legal(d, lxy)
        conv["legalization"] = {"seconds": time.perf_counter() - t0, "legal": chk["legal"],
                                "mean_displacement": lst["total_displacement"] / max(lst["cells"], 1),
                                "max_displacement": lst["max_displacement"], "check": chk}
        print(f"[bench] Abacus: {lst['cells']} cells in {conv['legalization']['seconds']:.2f}s, "
              f"{'legal' if chk['legal'] else 'NOT legal'}, mean displacement "
              f"{conv['legalization']['mean_displacement']:.3f}", file=sys.stderr, flush=True)
        if args.method == "bktrk" and not args.no_learned:
            from eplace_amd.controller import Controller
            ctrl = Controller.load()
            gp.write("positions", init_pos)
            device_sync()
            t0 = time.perf_counter()
            gp.initialization()
            rl = EplaceNesterovOpt(gp, method="learned", controller=ctrl, target_overflow=0.10,
                                   log=False).opt()
            device_sync()
            tl = time.perf_counter() - t0
            conv["learned"] = {"seconds_to_overflow_0.10": tl, "iterations": rl["iterations"],
                               "evaluations": rl["evaluations"], "backtracks": rl["backtracks"],
                               "controller_fallbacks": rl["controller_fallbacks"],
                               "final_overflow": rl["final_overflow"], "final_hpwl": rl["final_hpwl"],
                               "speedup_vs_plain": tc / tl, "hpwl_gap_vs_plain": rl["final_hpwl"] / rc["final_hpwl"] - 1}
            print(f"[bench] learned: {rl['iterations']} iterations in {tl:.3f}s ({tc / tl:.2f}x), "
                  f"overflow {rl['final_overflow']:.4f}, HPWL {rl['final_hpwl']:.6g} "
                  f"({100 * (rl['final_hpwl'] / rc['final_hpwl'] - 1):+.3f}%)", file=sys.stderr, flush=True)

    qp = None
    if not args.no_qp:
        from eplace_amd.qplace import quadratic_placement
        dq = synthetic(cells, nets, seed=args.seed, num_io=1000, util=0.7, name=f"synth_{cells}")
        quadratic_placement(dq, max_iterations=0)  # warm: code objects, allocator
        dq = synthetic(cells, nets, seed=args.seed, num_io=1000, util=0.7, name=f"synth_{cells}")
        t0 = time.perf_counter()
        sq = quadratic_placement(dq)
        tq = time.perf_counter() - t0
        qp = {"workload": "the same design, QPPlacer::quadraticPlacement defaults (from the core centre, "
                          "-IPiteCount 20, BiCGSTAB 100 iterations)",
              "seconds": tq, "iterations": sq["iterations"], "solver_iterations": sq["solver_iterations"],
              "seconds_per_outer_iteration": tq / sq["iterations"], "initial_hpwl": sq["initial_hpwl"],
              "final_hpwl": sq["final_hpwl"], "error_x": sq["error_x"], "error_y": sq["error_y"],
              "hpwl_log": sq["log"][:, 0].tolist()}
        print(f"[bench] QP: {sq['iterations']} outer iterations ({sq['solver_iterations']} BiCGSTAB) in {tq:.3f}s, "
              f"HPWL {sq['initial_hpwl']:.6g} -> {sq['final_hpwl']:.6g}", file=sys.stderr, flush=True)
        if not args.no_qp_cpu and not args.no_cpu:
            base = qp_reference_baseline(dq, 1)
            if base is not None:
                qp["cpu_baseline"] = base
                qp["speedup_per_outer_iteration"] = base["value"] / qp["seconds_per_outer_iteration"]
                print(f"[bench] QP reference: {base['value']:.2f}s per outer iteration "
                      f"({qp['speedup_per_outer_iteration']:.0f}x)", file=sys.stderr, flush=True)
        del dq

    copy_gbs = copy_peak_gbs()
    print(f"[bench] device copy {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    for v in [roof] + list(roof_all.values()):
        if v["bound"] == "hbm":
            v["copy_peak"] = copy_gbs
            v["frac_of_copy"] = v["achieved"] / copy_gbs

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(d, args.method, args.cpu_iters, init_pos)
        print(f"[bench] cpu baseline {cpu['value']:.4f} iters/s over {cpu['seconds']:.1f}s", file=sys.stderr,
              flush=True)

    out = {
        "metric": "GP time to overflow<=0.10 @1M cells, learned vs plain (HPWL +-0.5%); iters/s",
        "value": steps / dt,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": dt / steps * 1e3,
        "ms_per_step_with_phase_events": dt_events / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"synthetic std-cell design, {cells} cells, {nets} nets, {d.num_pins} pins, "
                               f"{info['num_fillers']} fillers, {info['bins_x']}x{info['bins_y']} bins, "
                               f"util 0.7, target density 1.0, seed {args.seed}",
                   "method": f"plain {args.method} Nesterov (device-resident)", "cells": cells,
                   "bins": [info["bins_x"], info["bins_y"]]},
        "roofline": roof,
        "roofline_phases": {k: {"bound": v["bound"], "achieved": v["achieved"], "unit": v["unit"],
                                "frac": v["frac"], "frac_of_copy": v.get("frac_of_copy"), "traffic": v["traffic"]}
                            for k, v in roof_all.items()},
        "phase_ms_per_evaluation": {k: v * 1e3 for k, v in ph.items()},
        "converge": conv,
        "qp": qp,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
