#!/usr/bin/env python3
"""Throughput of the MI355X Euler step on the 2D 4096^2 split MUSCL-Hancock
+ van Leer + HLLC workload (FP64), as cell-updates/s and HBM roofline.

A step is one full time step of the simulator loop
(simulator/simulator_impl.hpp:235-252): dt from the max wavespeed (fused into
the previous step's last sweep, computed on the device), the x sweep and the
y sweep.  Inputs are resident in HBM before the timed region starts.

The headline runs the FP64 step in the fast numerics (FP64 arithmetic
throughout, held to 1e-12 relative L-inf of the reference's results by the
GPU tests); the bit-exact FP64 numerics run beside it as the variant
hllc_f64_exact (--numerics exact makes it the headline).

Prints ONE JSON line.  Only the cpu_baseline leg loads the CPU oracle.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0        # MI355X published HBM3E bandwidth
FP64_PEAK_TFLOPS = 78.6

# Launch shapes (x lanes, x rows, y lanes, y rows) of the headline step per
# FP64 numerics.  The kernel trace and the HBM counters in
# profiles/sweeps_4096_f64_hllc.md were taken with these shapes (the
# autotuner's choices on this grid; --autotune-headline times the candidates
# instead).
HEADLINE_TUNING = {"fast": (128, 90, 64, 86), "exact": (128, 18, 64, 43)}

# Per FP64 numerics, at HEADLINE_TUNING, from profiles/sweeps_4096_f64_hllc.md
# (profile-derived, not measured by this run; the line says so in
# roofline.traffic_source):
#   traffic_ratio: HBM bytes / algorithmic bytes per launch (rocprofv3
#     FETCH_SIZE and WRITE_SIZE, calibrated against a plain device copy);
#   valu_per_cell: VALU instructions per interior cell (SQ_INSTS_VALU x 64
#     lanes / cells);
#   f64_flop_per_cell: FP64 operations per interior cell
#     (SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes / cells).
PROFILE = {
    "fast": {"traffic_ratio": {"x": 1.228, "y": 1.316},
             "valu_per_cell": {"x": 396.2, "y": 390.2},
             "f64_flop_per_cell": {"x": 283.8, "y": 277.4}},
    "exact": {"traffic_ratio": {"x": 1.403, "y": 1.386},
              "valu_per_cell": {"x": 519.9, "y": 514.9},
              "f64_flop_per_cell": {"x": 436.9, "y": 440.9}},
}

# VALU issue peak of the MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64
# instruction per SIMD every 4 cycles = 64 lanes x 614.4 G = 39.3 T lane
# instructions/s (FP64 FMA: 16 lanes per SIMD per clock, 78.6 TFLOP/s).
VALU_PEAK_LANE_INST_S = 256 * 4 * 2.4e9 / 4 * 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-variants", action="store_true", help="skip FORCE / FP32 lines")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the built-in launch shapes instead of timing candidates first")
    ap.add_argument("--autotune-headline", action="store_true",
                    help="time the launch-shape candidates for the headline step too "
                         "(default: the profiled shapes; `traffic` is then null)")
    ap.add_argument("--x-block", type=int, default=0)
    ap.add_argument("--x-rows", type=int, default=0)
    ap.add_argument("--y-block", type=int, default=0)
    ap.add_argument("--y-rows", type=int, default=0)
    ap.add_argument("--mode", choices=["bench", "sweeps", "ghost", "batch", "adjoint", "io"],
                    default="bench",
                    help="sweeps: a plain device copy of one grid (HBM calibration for "
                         "counter runs) followed by --steps x sweeps and y sweeps")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64",
                    help="precision of --mode sweeps and --mode ghost")
    ap.add_argument("--flux", choices=["hllc", "force"], default="hllc",
                    help="flux of --mode sweeps")
    ap.add_argument("--numerics", choices=["exact", "fast"], default="fast",
                    help="FP64 numerics of the headline step and of --mode sweeps (fast: "
                         "within 1e-12 relative L-inf of the reference; exact: its bits)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU decomposition is a design, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from fluidity_amd import euler as E
    from fluidity_amd import native as N
    from fluidity_amd import workloads as W

    if args.mode == "adjoint":
        stream = torch.cuda.Stream(device=torch.device("cuda:0"))
        print(json.dumps(adjoint_case(torch, N, W, stream, steps=args.steps)), flush=True)
        return
    if args.mode == "batch":
        stream = torch.cuda.Stream(device=torch.device("cuda:0"))
        print(json.dumps(batched_case(torch, N, W, stream, args.steps, numerics=args.numerics)),
              flush=True)
        return
    if args.mode == "io":
        stream = torch.cuda.Stream(device=torch.device("cuda:0"))
        print(json.dumps(io_case(torch, N, W, stream)), flush=True)
        return
    if args.mode == "ghost":
        stream = torch.cuda.Stream(device=torch.device("cuda:0"))
        print(json.dumps(ghost_fluid_case(torch, N, W, stream, args.steps, args.warmup,
                                          dtype=torch.float32 if args.dtype == "f32" else None,
                                          numerics=args.numerics)),
              flush=True)
        return
    n = args.n
    wl = W.riemann_2d(n=n, seed=1001)
    if args.mode == "sweeps":
        return sweeps_mode(args, wl, torch, E, N, W)
    dh = wl.dh
    cfl = 0.9
    cells = n * n
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)

    def make_solver(flux, layout=N.FL_LAYOUT_SOA, numerics=N.FL_NUMERICS_EXACT):
        s = E.SplitSolver(E.Scheme(flux=flux, numerics=numerics), layout=layout)
        s.set_tuning(x_block=args.x_block, x_rows=args.x_rows, y_block=args.y_block,
                     y_rows=args.y_rows)
        return s

    tuned = {}

    U_soa = W.aos_to_soa(wl.U)

    def run_case(flux, dtype, steps, warmup, numerics=N.FL_NUMERICS_EXACT):
        solver = make_solver(flux, numerics=numerics)
        with torch.cuda.stream(stream):
            a = torch.from_numpy(U_soa).to(device=dev, dtype=dtype)
            w = torch.empty_like(a)
        stream.synchronize()
        explicit = args.x_block or args.x_rows or args.y_block or args.y_rows
        headline = flux == N.FL_FLUX_HLLC and dtype == torch.float64
        if headline and not explicit and not args.autotune_headline:
            shape = HEADLINE_TUNING["fast" if numerics == N.FL_NUMERICS_FAST else "exact"]
            xb, xr, yb, yr = shape
            solver.set_tuning(x_block=xb, x_rows=xr, y_block=yb, y_rows=yr)
            tuned[(flux, str(dtype), numerics)] = list(shape)
        elif args.autotune and not explicit:
            t = solver.autotune(a, w, reps=2, stream=stream)
            tuned[(flux, str(dtype), numerics)] = [t.x_block, t.x_rows, t.y_block, t.y_rows]
        if warmup:
            solver.simulate(a, w, cfl, dh, max_iters=warmup, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = solver.simulate(a, w, cfl, dh, max_iters=steps, stream=stream)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        assert r.iters == steps, (r.iters, steps)
        return solver, a, w, el, r

    # ---- headline: FP64 HLLC
    head_num = N.FL_NUMERICS_FAST if args.numerics == "fast" else N.FL_NUMERICS_EXACT
    solver, a, w, elapsed, res = run_case(N.FL_FLUX_HLLC, torch.float64, args.steps, args.warmup,
                                          head_num)
    value = args.steps * cells / elapsed
    ms_per_step = elapsed * 1e3 / args.steps

    # ---- per-sweep kernel time with HIP events on the launch stream
    dtdh = res.last_dt / dh
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    reps = max(5, min(args.steps, 20))
    with torch.cuda.stream(stream):
        ev[0].record(stream)
        for _ in range(reps):
            solver.sweep(a, w, 0, dtdh, stream=stream)
        ev[1].record(stream)
        for _ in range(reps):
            solver.sweep(w, a, 1, dtdh, stream=stream)
        ev[2].record(stream)
    stream.synchronize()
    t_x = ev[0].elapsed_time(ev[1]) / reps * 1e-3
    t_y = ev[1].elapsed_time(ev[2]) / reps * 1e-3
    bytes_per_sweep = cells * 64.0            # read 32 B + write 32 B per cell (FP64, 2D)
    kind = "fast" if head_num == N.FL_NUMERICS_FAST else "exact"
    dom_name, dom_t, dom_k = (f"sweep_x_kernel<double,2,HLLC,{kind},van Leer>", t_x, "x") \
        if t_x >= t_y else (f"sweep_y_kernel<double,HLLC,{kind},van Leer>", t_y, "y")
    achieved = bytes_per_sweep / dom_t / 1e9
    profiled = (n == 4096 and tuned.get((N.FL_FLUX_HLLC, str(torch.float64), head_num))
                == list(HEADLINE_TUNING[kind]))
    prof = PROFILE[kind]
    traffic = round(prof["traffic_ratio"][dom_k] * bytes_per_sweep) if profiled else None
    copy_gbs = hbm_copy_ceiling(torch, stream)
    valu_rate = prof["valu_per_cell"][dom_k] * cells / dom_t
    roofline = {
        # The sweeps issue FP64 VALU work at about half the issue peak and move
        # HBM bytes at about a third of its peak: VALU issue bounds them.
        "bound": "valu",
        "kernel": dom_name,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE / algorithmic bytes at these "
                           "launch shapes (profiles/sweeps_4096_f64_hllc.md), times this "
                           "launch's algorithmic bytes; not counted in this run")
        if traffic is not None else None,
        "algorithmic_bytes": int(bytes_per_sweep),
        "hbm_copy_GBps": round(copy_gbs, 1),
        "frac_of_copy": round(achieved / copy_gbs, 4),
        "valu_insts_per_cell": prof["valu_per_cell"][dom_k] if profiled else None,
        "valu_achieved_Tinst_s": round(valu_rate / 1e12, 2) if profiled else None,
        "valu_peak_Tinst_s": round(VALU_PEAK_LANE_INST_S / 1e12, 2),
        "valu_frac": round(valu_rate / VALU_PEAK_LANE_INST_S, 4) if profiled else None,
        "fp64_TFLOPs": (round(prof["f64_flop_per_cell"][dom_k] * cells / dom_t / 1e12, 2)
                        if profiled else None),
        "fp64_peak_TFLOPs": FP64_PEAK_TFLOPS,
        "numerics": kind,
        "tuning": tuned.get((N.FL_FLUX_HLLC, str(torch.float64), head_num)),
        "x_sweep_us": round(t_x * 1e6, 1),
        "y_sweep_us": round(t_y * 1e6, 1),
        "step_bytes_per_cell": 128,
        "step_frac": round(value * 128 / 1e9 / HBM_PEAK_GBS, 4),
        "step_frac_of_copy": round(value * 128 / 1e9 / copy_gbs, 4),
    }
    del a, w, solver
    torch.cuda.empty_cache()

    variants = {}
    # the FP64 step in the other numerics: the bit-exact one beside a fast
    # headline (or the fast one beside an exact headline)
    other = N.FL_NUMERICS_EXACT if head_num == N.FL_NUMERICS_FAST else N.FL_NUMERICS_FAST
    exact_value = value if head_num == N.FL_NUMERICS_EXACT else None
    if not args.no_variants:
        vsteps = max(10, args.steps // 2)
        for name, flux, dt, bpc, num in [
                ("hllc_f64_exact" if other == N.FL_NUMERICS_EXACT else "hllc_f64_fast",
                 N.FL_FLUX_HLLC, torch.float64, 128, other),
                ("force_f64", N.FL_FLUX_FORCE, torch.float64, 128, N.FL_NUMERICS_EXACT),
                ("force_f64_fast", N.FL_FLUX_FORCE, torch.float64, 128, N.FL_NUMERICS_FAST),
                ("hllc_f32", N.FL_FLUX_HLLC, torch.float32, 64, N.FL_NUMERICS_EXACT)]:
            s2, a2, w2, el2, _ = run_case(flux, dt, vsteps, max(2, args.warmup // 2), num)
            v = vsteps * cells / el2
            if name == "hllc_f64_exact":
                exact_value = v
            variants[name] = {"value": round(v, 1), "unit": "cell-updates/s",
                              "ms_per_step": round(el2 * 1e3 / vsteps, 4),
                              "hbm_frac": round(v * bpc / 1e9 / HBM_PEAK_GBS, 4),
                              "vs_f64_hllc": round(v / value, 3),
                              "tuning": tuned.get((flux, str(dt), num))}
            del s2, a2, w2
            torch.cuda.empty_cache()
        variants["hllc_f32"]["vs_f64_exact"] = round(variants["hllc_f32"]["value"] / exact_value, 3)

        variants["riemann_1d_f64"] = riemann_1d_case(torch, N, W, stream)
        # the two-material step in the headline's FP64 numerics, the other
        # one beside it
        other = "exact" if args.numerics == "fast" else "fast"
        variants["ghost_fluid_f64"] = ghost_fluid_case(torch, N, W, stream, steps=10, warmup=5,
                                                       numerics=args.numerics)
        variants[f"ghost_fluid_f64_{other}"] = ghost_fluid_case(torch, N, W, stream, steps=10,
                                                                warmup=5, numerics=other)
        variants["ghost_fluid_f32"] = ghost_fluid_case(torch, N, W, stream, steps=10, warmup=5,
                                                       dtype=torch.float32)
        variants["adjoint_f64"] = adjoint_case(torch, N, W, stream)
        variants["io_f64"] = io_case(torch, N, W, stream)
        bf = batched_case(torch, N, W, stream, steps=20, numerics=args.numerics)
        for k, v in bf.items():
            # against the headline FP64 step and against the bit-exact one
            v["vs_f64_hllc"] = round(v["value"] / value, 3)
            v["vs_f64_exact"] = round(v["value"] / exact_value, 3)
            variants[k] = v

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, cfl, args.cpu_steps)
        if not args.no_variants:
            cpu["riemann_1d_f64"] = cpu_riemann_1d(W)
            cpu["ghost_fluid_f64"] = cpu_ghost_fluid(W)

    line = {
        "metric": "cell-updates/s, 2D 4096^2 split MUSCL-Hancock+HLLC FP64",
        "value": round(value, 1),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded four-quadrant Riemann data, seed 1001)",
        "config": {"workload": f"2D Euler {n}x{n}, gamma 1.4, MUSCL-Hancock + van Leer + HLLC, "
                               "dimensionally split x then y, transmissive, CFL 0.9",
                   "nx": n, "ny": n, "layout": "SoA", "parallelism": "single GPU",
                   "numerics": ("FP64 fast: within 1e-12 relative L-inf of the reference"
                                if kind == "fast" else "FP64 exact: the reference's bits")},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "variants": variants,
    }
    print(json.dumps(line), flush=True)


def hbm_copy_ceiling(torch, stream, nbytes=1 << 30, reps=10):
    """The box's measured device-copy bandwidth: a 1 GiB device-to-device
    copy (read 1 GiB + write 1 GiB) timed with HIP events on the launch
    stream, best of `reps`."""
    dev = torch.device("cuda:0")
    with torch.cuda.stream(stream):
        src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        best = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            dst.copy_(src)
            e1.record(stream)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / best / 1e9


def riemann_1d_case(torch, N, W, stream, runs=3):
    """The 1D Riemann workload (4096 cells, seed 1000, HLLC + van Leer, CFL
    0.9) run until the fastest wave has travelled from x = 0.5 to 0.9, as a
    whole: the device step loop with graph-replayed steps, timed over `runs`
    runs after one untimed run.  On 4096 cells a step is a few microseconds
    of launches, so this measures the step loop's launch overhead more than
    the sweep."""
    from fluidity_amd import euler as E
    r1 = W.riemann_1d()
    dev = torch.device("cuda:0")
    solver = E.SplitSolver(E.Scheme())
    with torch.cuda.stream(stream):
        a0 = torch.from_numpy(r1.U).to(dev)
        a = a0.clone()
        w = torch.empty_like(a)
    stream.synchronize()
    solver.simulate(a, w, 0.9, r1.dh, sim_time=r1.t_end, stream=stream)   # untimed
    times, iters = [], 0
    for _ in range(runs):
        with torch.cuda.stream(stream):
            a.copy_(a0)
        stream.synchronize()
        t0 = time.perf_counter()
        res = solver.simulate(a, w, 0.9, r1.dh, sim_time=r1.t_end, stream=stream)
        stream.synchronize()
        times.append(time.perf_counter() - t0)
        iters = res.iters
    el = min(times)
    n = r1.U.shape[0]
    return {"value": round(iters * n / el, 1), "unit": "cell-updates/s",
            "run_ms": round(el * 1e3, 3), "steps": iters, "us_per_step": round(el * 1e6 / iters, 2),
            "workload": f"1D Euler {n} cells, seed 1000, HLLC + van Leer, CFL 0.9, until the "
                        "fastest wave has travelled 0.4 (x = 0.5 -> 0.9), FP64"}


def cpu_riemann_1d(W):
    """The same 1D run on the CPU oracle, single-threaded, as the reference's
    CPU path runs it."""
    from oracle import oracle as O
    r1 = W.riemann_1d()
    t0 = time.perf_counter()
    _, _, it, _ = O.simulate(r1.U, O.scheme(flux=O.FLUX_HLLC), 0.9, r1.dh, sim_time=r1.t_end,
                             max_iters=1000000, nthreads=1)
    el = time.perf_counter() - t0
    n = r1.U.shape[0]
    return {"value": round(it * n / el, 1), "unit": "cell-updates/s", "cores": 1,
            "kind": "port", "run_ms": round(el * 1e3, 1), "steps": it,
            "sample": f"the whole 1D run ({it} steps of {n} cells)"}


def cpu_ghost_fluid(W, steps=5, nx=2048, ny=1024):
    """The two-material shock-bubble step on the CPU oracle (OpenMP), a
    bounded sample: `steps` steps from the initial state, one of them with the
    reinitialisation (every 5 steps)."""
    from oracle import oracle as O
    sb = W.shock_bubble(nx, ny, seed=1002)
    threads = min(os.cpu_count() or 1, 16)
    s0, s1 = O.scheme(gamma=sb.gamma0), O.scheme(gamma=sb.gamma1)
    t0 = time.perf_counter()
    R = O.mm_simulate(sb.U0, sb.U1, sb.phi0, sb.phi1, s0, s1, 0.5, sb.dh, max_iters=steps,
                      reinit_every=5, nthreads=threads)
    el = time.perf_counter() - t0
    return {"value": round(R[5] * nx * ny / el, 1), "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "ms_per_step": round(el * 1e3 / R[5], 1),
            "sample": f"{R[5]} two-material steps of the {nx}x{ny} shock-bubble workload "
                      f"(one with reinitialisation), {el:.1f} s"}


def adjoint_case(torch, N, W, stream, steps=200, every=20, n=1024):
    """Gradient of the density misfit after 200 steps of the 1024^2 periodic
    bump workload (seed 1004, target seed 1005) with respect to the initial
    state and gamma: the checkpointed forward run (a checkpoint every 20
    steps) and the hand-written adjoint with per-segment recompute, FP64.
    `value` counts adjoint cell-steps (cells x steps) per second of the
    backward pass, which includes the recompute."""
    from fluidity_amd import adjoint as AJ
    from fluidity_amd import euler as E
    prob = W.adjoint_problem(n)
    adj = AJ.Adjoint(E.Scheme(gamma=prob.gamma, bc=(2, 2, 0)), 0.9, prob.dh)
    dev = torch.device("cuda:0")
    with torch.cuda.stream(stream):
        U = torch.from_numpy(np.ascontiguousarray(np.moveaxis(prob.U, -1, 0))).to(dev)
        tgt = torch.from_numpy(prob.target).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run = adj.forward(U, steps, every, stream=stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    with torch.cuda.stream(stream):
        grad = torch.zeros_like(U)
        grad[0] = 2.0 * prob.dh * prob.dh * (U[0] - tgt)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    dgamma = adj.backward(grad, run, stream=stream)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    cells = n * n
    return {"value": round(steps * cells / (t3 - t2), 1), "unit": "adjoint cell-steps/s",
            "forward_ms": round((t1 - t0) * 1e3, 2), "backward_ms": round((t3 - t2) * 1e3, 2),
            "steps": steps, "checkpoint_every": every,
            "checkpoint_bytes": int(run["checkpoints"].numel()),
            "dloss_dgamma": dgamma,
            "workload": f"2D Euler {n}x{n} periodic, 16 density bumps (seed 1004), u=v=0, p=1, "
                        "HLLC + van Leer, CFL 0.9, FP64"}


def batched_case(torch, N, W, stream, steps, numerics="exact"):
    """The 512-problem batch (seed 1003: 384 1D problems of 256..65536 cells,
    128 2D problems of 64^2..1024^2, per-problem gamma and dt), FP64 in the
    given numerics (batched_f64), FP64 in the other one beside it and FP32
    bulk, tile shapes from the autotuner, steps replayed from a hipGraph.  A
    cell-update is one cell of one problem advanced one step.  The FP32 and
    fast FP64 lines carry their per-problem density error against the exact
    FP64 batch, which equals the oracle run on each problem alone bit for
    bit."""
    from fluidity_amd import batch as BT
    specs = W.batched(seed=1003)
    cells = sum(int(np.prod(sp.n)) for sp in specs)
    # the exact FP64 batch first: the reference of the others
    runs = [("batched_f64" if num == numerics else f"batched_f64_{num}", torch.float64, num)
            for num in ("exact", "fast")] + [("batched_f32", torch.float32, "exact")]
    out, ref = {}, None
    for name, dt, num in runs:
        probs = BT.problems_from_workload(
            specs, dt, N.FL_LAYOUT_SOA, max_iters=steps + 2,
            numerics=N.FL_NUMERICS_FAST if num == "fast" else N.FL_NUMERICS_EXACT)
        b = BT.Batch(probs, stream=stream)
        t = b.autotune(reps=2, stream=stream)
        b.step(2, graph=True, stream=stream)          # warm-up (captures the graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.step(steps, graph=True, stream=stream)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res = b.finish(allow_failed=True, stream=stream)
        # problems the scheme cannot run stop after their NaN step: count the
        # cell-updates actually done in the timed steps
        done = sum(int(np.prod(sp.n)) * max(0, r.iters - 2) for sp, r in zip(specs, res))
        ok = [bool(np.isfinite(r.max_wavespeed) and r.max_wavespeed > 0) for r in res]
        line = {"value": round(done / el, 1), "unit": "cell-updates/s",
                "ms_per_step": round(el * 1e3 / steps, 4), "problems": len(specs),
                "failed_problems": ok.count(False), "cells": cells,
                "tuning": [t.x_block, t.x_rows, t.y_block, t.y_rows]}
        if dt == torch.float64:
            line["numerics"] = num
        if ref is None:
            ref = [(b.state(i)[0].clone(), ok[i]) for i in range(len(specs))]
        else:
            pairs = [(b.state(i)[0].double(), ref[i][0])
                     for i in range(len(specs)) if ref[i][1] and ok[i]]
            errs = [float((d - r).abs().sum() / r.abs().sum()) for d, r in pairs]
            line["density_rel_l1_max"] = float("%.3g" % max(errs))
            line["density_rel_l1_median"] = float("%.3g" % float(np.median(errs)))
            line["density_rel_linf_max"] = float("%.3g" % max(
                float((d - r).abs().max() / r.abs().max()) for d, r in pairs))
            line["error_problems"] = len(errs)
        out[name] = line
        del b, probs
        torch.cuda.empty_cache()
    return out


def io_case(torch, N, W, stream, n=4096, reps=20, text_n=512):
    """The device halves of the initial-condition and output paths on the
    4096^2 FP64 grid: conservative -> primitive planes with internal energy
    (read 32 B, write 40 B per cell) and primitive planes -> conservative
    (read 32 B, write 32 B per cell), each timed with HIP events on the launch
    stream over `reps` launches; then the reference's ASCII result file of a
    512^2 grid written to a temporary directory (host formatting)."""
    import tempfile
    from fluidity_amd import simulator as SM
    from fluidity_amd.euler import Scheme
    wl = W.riemann_2d(n=n, seed=1001)
    dev = torch.device("cuda:0")
    sch = Scheme()
    with torch.cuda.stream(stream):
        U = torch.from_numpy(W.aos_to_soa(wl.U)).to(dev)
        prim = SM.primitive_planes(U, sch, N.FL_LAYOUT_SOA, stream=stream)
        out = torch.empty_like(U)
    stream.synchronize()
    cells = n * n
    res = {"workload": f"2D {n}x{n} FP64 SoA (seed 1001); ASCII file of a {text_n}^2 grid"}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    with torch.cuda.stream(stream):  # the planes are allocated on the launch stream
        prim4 = prim[:4].contiguous()
        SM.from_primitive(prim4, sch, out, N.FL_LAYOUT_SOA, stream=stream)
        ev[0].record(stream)
        for _ in range(reps):
            SM.primitive_planes(U, sch, N.FL_LAYOUT_SOA, stream=stream)
        ev[1].record(stream)
        for _ in range(reps):
            SM.from_primitive(prim4, sch, out, N.FL_LAYOUT_SOA, stream=stream)
        ev[2].record(stream)
    stream.synchronize()
    t_prim = ev[0].elapsed_time(ev[1]) / reps * 1e-3
    t_cons = ev[1].elapsed_time(ev[2]) / reps * 1e-3
    res["primitive_planes_us"] = round(t_prim * 1e6, 1)
    res["primitive_planes_GBps"] = round(72 * cells / t_prim / 1e9, 1)
    res["primitive_planes_hbm_frac"] = round(72 * cells / t_prim / 1e9 / HBM_PEAK_GBS, 4)
    res["from_primitive_us"] = round(t_cons * 1e6, 1)
    res["from_primitive_GBps"] = round(64 * cells / t_cons / 1e9, 1)
    res["from_primitive_hbm_frac"] = round(64 * cells / t_cons / 1e9 / HBM_PEAK_GBS, 4)
    small = W.riemann_2d(n=text_n, seed=1001)
    sim = SM.Simulator(dimensions=2)
    sim.configure_resolution(1.0 / text_n)
    sim.configure_dimension(0, 0.0, 1.0)
    sim.configure_dimension(1, 0.0, 1.0)
    sim.state.copy_(torch.from_numpy(W.aos_to_soa(small.U)).to(dev))
    torch.cuda.synchronize()
    with tempfile.TemporaryDirectory() as d:
        t0 = time.perf_counter()
        sim.write_results(os.path.join(d, "io"))
        el = time.perf_counter() - t0
        size = os.path.getsize(os.path.join(d, "io.dat"))
    res["write_results_cells_per_s"] = round(text_n * text_n / el, 1)
    res["write_results_MBps"] = round(size / el / 1e6, 1)
    del U, prim, prim4, out
    torch.cuda.empty_cache()
    return res


def ghost_fluid_case(torch, N, W, stream, steps, warmup, nx=2048, ny=1024, dtype=None,
                     numerics="exact"):
    """Two-material ghost-fluid steps on the seeded shock-bubble workload
    (ideal-gas bubble, level sets reinitialised every 5 steps in a 6-cell
    band).  A cell-update is one cell of the domain advanced one full
    multimaterial step (both materials' sweeps, ghost states, level-set
    advection, projection and the share of reinitialisation).  With 5 warm-up
    steps the first reinitialisation falls in the warm-up call, so the timed
    steps are the workload's steady state (level sets clamped outside the
    band, as after every reinitialisation).  dtype float32: FP32 bulk states
    (FP64 level sets, dt and ghost-fluid arithmetic).  numerics "fast": the
    FP64 sweeps and the level sets' WENO right-hand side in the fast numerics
    (within 1e-12 of the exact loop)."""
    from fluidity_amd import euler as E
    from fluidity_amd import levelset as L
    sb = W.shock_bubble(nx, ny, seed=1002)
    dev = torch.device("cuda:0")
    sim = L.MultimaterialSimulator(layout=N.FL_LAYOUT_SOA)
    sdt = dtype or torch.float64
    with torch.cuda.stream(stream):
        U0 = torch.from_numpy(W.aos_to_soa(sb.U0)).to(dev, dtype=sdt)
        U1 = torch.from_numpy(W.aos_to_soa(sb.U1)).to(dev, dtype=sdt)
        p0 = torch.from_numpy(sb.phi0).to(dev)
        p1 = torch.from_numpy(sb.phi1).to(dev)
    stream.synchronize()
    num = N.FL_NUMERICS_FAST if numerics == "fast" else N.FL_NUMERICS_EXACT
    sim.add_material(U0, p0, E.Scheme(gamma=sb.gamma0, numerics=num))
    sim.add_material(U1, p1, E.Scheme(gamma=sb.gamma1, numerics=num))
    kw = dict(band=6, extrap_sweeps=8, reinit_every=5, stream=stream)
    sim.simulate_mm(0.5, sb.dh, max_iters=warmup, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = sim.simulate_mm(0.5, sb.dh, max_iters=steps, **kw)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    assert r.iters == steps
    cells = nx * ny
    return {"value": round(steps * cells / el, 1), "unit": "cell-updates/s",
            "ms_per_step": round(el * 1e3 / steps, 4), "steps": steps,
            "warmup": warmup, "numerics": numerics,
            "workload": f"shock-bubble {nx}x{ny} {'FP32 bulk' if dtype else 'FP64'}, 2 materials, "
                        f"Mach {sb.mach:.3f}, reinit every 5 steps, band 6, SoA"}


def sweeps_mode(args, wl, torch, E, N, W):
    """Profiling driver: one device copy of a grid-sized buffer (the plain copy
    kernel whose counters calibrate FETCH_SIZE / WRITE_SIZE), then --steps x
    sweeps and --steps y sweeps of the 4096^2 workload."""
    n = args.n
    dt = torch.float64 if args.dtype == "f64" else torch.float32
    flux = N.FL_FLUX_HLLC if args.flux == "hllc" else N.FL_FLUX_FORCE
    num = N.FL_NUMERICS_FAST if args.numerics == "fast" else N.FL_NUMERICS_EXACT
    solver = E.SplitSolver(E.Scheme(flux=flux, numerics=num), layout=N.FL_LAYOUT_SOA)
    if args.x_block or args.x_rows or args.y_block or args.y_rows:
        solver.set_tuning(x_block=args.x_block, x_rows=args.x_rows, y_block=args.y_block,
                          y_rows=args.y_rows)
    else:
        xb, xr, yb, yr = HEADLINE_TUNING[args.numerics]
        solver.set_tuning(x_block=xb, x_rows=xr, y_block=yb, y_rows=yr)
    a = torch.from_numpy(W.aos_to_soa(wl.U)).to(device="cuda:0", dtype=dt)
    w = torch.empty_like(a)
    dtdh = 0.9 / solver.max_wavespeed(a)
    w.copy_(a)                        # calibration copy: reads and writes one grid
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record()
    for _ in range(args.steps):
        solver.sweep(a, w, 0, dtdh)
    ev[1].record()
    for _ in range(args.steps):
        solver.sweep(w, a, 1, dtdh)
    ev[2].record()
    torch.cuda.synchronize()
    tx = ev[0].elapsed_time(ev[1]) / args.steps
    ty = ev[1].elapsed_time(ev[2]) / args.steps
    grid_bytes = a.numel() * a.element_size()
    print(json.dumps({"mode": "sweeps", "dtype": args.dtype, "flux": args.flux, "n": n,
                      "numerics": args.numerics,
                      "grid_bytes": grid_bytes, "x_sweep_ms": round(tx, 4),
                      "y_sweep_ms": round(ty, 4),
                      "x_GBps": round(2 * grid_bytes / tx / 1e6, 1),
                      "y_GBps": round(2 * grid_bytes / ty / 1e6, 1)}), flush=True)


def cpu_baseline(wl, cfl, steps):
    """The CPU oracle (a restatement of the reference path, OpenMP over rows /
    columns) on the same 4096^2 grid for a bounded number of steps."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    s = O.scheme(flux=O.FLUX_HLLC)
    t0 = time.perf_counter()
    _, _, it, _ = O.simulate(wl.U, s, cfl, wl.dh, max_iters=steps, nthreads=threads)
    el = time.perf_counter() - t0
    cells = wl.U.shape[0] * wl.U.shape[1]
    return {"value": round(it * cells / el, 1), "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{it} full step(s) of the same {wl.U.shape[1]}x{wl.U.shape[0]} FP64 "
                      f"HLLC workload, {el:.1f} s"}


if __name__ == "__main__":
    main()
