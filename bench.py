#!/usr/bin/env python3
"""Benchmark of the FVM time step on one MI355X.

Workload (the configuration the headline metric is quoted on): viscous
Navier-Stokes on a 2048x1024 grid, FP32 state, Mach 0.3, Re 200 on a 64-cell
cylinder at (512,512), Sutherland viscosity, no-slip adiabatic wall, HLLC +
van Leer MUSCL, SSP-RK2 with Strang viscous half steps, CFL 0.4 / 0.4 with the
CFL dt recomputed on the device every step, +-1e-3 seeded transverse-velocity
noise.  A step is one full time step of every cell; the step sequence is
replayed from a captured HIP graph with the state resident in HBM.

Prints ONE JSON line.  `value` = cell-updates/s over exactly --steps timed
steps, bracketed by device synchronisation.  Before them run untimed steps
for --settle seconds (the clocks ramp up; the count is `clock_settle_steps`)
and then --warmup untimed steps.
`roofline` = the dominant kernel's (the one moving the most algorithmic bytes
per step) algorithmic HBM bytes per launch over its average launch duration,
measured with HIP events on the solver's stream in a timing pass right after
the timed region; `roofline.per_kernel` gives every kernel's figures.  `cpu_baseline` = the CPU oracle
(FP32, OpenMP) on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "fvm-airidynamic-simulator_amd"))

HBM_PEAK = 8.0e12  # bytes/s, MI355X
# Vector-ALU issue peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one FP32 lane
# operation per lane per cycle (no packed math), i.e. a wave64 instruction
# every 4 cycles per SIMD.
VALU_PEAK = 256 * 4 * 16 * 2.4e9
# Algorithmic HBM bytes per cell for each kernel of one step (FP32 state of
# 5 fields = 20 B, cell tag 1 B): what the kernel must read and write once.
KERNEL_BYTES_PER_CELL = {
    "diffuse_half1": 20 + 1 + 16,        # read U^n, tag; write U_d (rho is U^n's plane)
    "stage1": 20 + 1 + 20,               # read U_d (or U^n), tag; write U*
    "stage1_fused": 20 + 1 + 16 + 20,    # read U^n, tag; write U_d and U* (fused schedule)
    "stage2": 20 + 20 + 1 + 20,          # read U*, U_d (blend), tag; write blend
    "diffuse_half2": 20 + 1 + 20,        # read blend, tag; write U^{n+1}
}
B_ALG_PER_CELL_UPDATE = 102  # 5 x 20 B of carried state + 2 x 1 B tag per step


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--settle", type=float, default=0.5,
                    help="seconds of untimed steps before the warm-up steps, for the clocks to ramp up")
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=200, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--batch-cases", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=2000, help="steps of every case (the mix specifies 2000)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=8.0,
                    help="requests/s of the served stream (8/s offers ~8.3e9 cell-updates/s, about 75%% of capacity)")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-step-scale", type=float, default=1.0, help="scales the 500-5000 step counts")
    ap.add_argument("--serve-time-limit", type=float, default=None,
                    help="stop the served stream after this many seconds (default 3x the window)")
    ap.add_argument("--overload-rate", type=float, default=50.0, help="requests/s of the short overload stream")
    ap.add_argument("--overload-seconds", type=float, default=4.0, help="0: no overload stream")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-other", action="store_true", help="skip the 256x128 Euler case and the CPU mix subset")
    ap.add_argument("--direct", action="store_true",
                    help="launch every kernel directly instead of replaying captured graphs (kernel tracing)")
    return ap.parse_args()


def counters_from_profiles(nx, ny):
    """Per-launch counters of the step kernels from the committed rocprof PMC
    summary: HBM bytes (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel
    factor) and VALU instructions (SQ_INSTS_VALU).  The summary covers the
    2048x1024 workload only; other grids get {}."""
    path = os.path.join(REPO, "profiles", "stage_kernels_2048x1024_pmc.json")
    if (nx, ny) != (2048, 1024) or not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            return json.load(f)
    except Exception:
        return {}


def cpu_baseline(p, nsteps):
    """The FP32 oracle with OpenMP on this host, nsteps of the same workload."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    threads = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1)
    q = p.copy()
    q.dt_mode = 1
    o = O.Oracle(q, fp64=False)
    t0 = time.perf_counter()
    o.step(nsteps, q)
    dt = time.perf_counter() - t0
    o.close()
    rate = p.nx * p.ny * nsteps / dt
    return {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} steps of the {p.nx}x{p.ny} viscous workload, FP32 oracle, OpenMP, {dt:.1f} s"}


def cpu_rate(p, nsteps, threads):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    O.set_threads(threads)
    q = p.copy()
    o = O.Oracle(q, fp64=False)
    t0 = time.perf_counter()
    o.step(nsteps, q)
    dt = time.perf_counter() - t0
    o.close()
    return p.nx * p.ny * nsteps / dt, dt


def other_configs(with_cpu):
    """The other workloads: the 256x128 Mach-3 Euler case (launch-latency
    bound on the GPU) in full, 1000 steps, on the GPU and on the CPU oracle;
    and a 32-case x 200-step subset of the batch mix on the CPU oracle."""
    import fvm_amd
    from fvm_amd import params as P
    out = {}
    p = P.euler_bow_shock()
    p.dt_mode = P.DT_DEVICE_CFL
    s = fvm_amd.CFDSolver(0)
    s.initialize(p)
    q = p.copy()
    s.advance(q, 100)
    s.synchronize()
    t0 = time.perf_counter()
    s.advance(q, 1000)
    s.synchronize()
    dt = time.perf_counter() - t0
    s.close()
    out["euler_256x128"] = {"gpu_cell_updates_per_s": p.nx * p.ny * 1000 / dt, "gpu_us_per_step": dt / 1000 * 1e6}
    # the shape of the reference's README screenshot run (SI Mach-3 circle r = 0.5 m
    # in a 10 x 5 m domain on 1486 x 1828 cells, inviscid, CFL 0.5): context only
    r = P.default_params()
    r.nx, r.ny = 1486, 1828
    P.compute_derived(r)
    r.dt_mode = P.DT_DEVICE_CFL
    s = fvm_amd.CFDSolver(0)
    s.initialize(r)
    q = r.copy()
    s.advance(q, 20)
    s.synchronize()
    t0 = time.perf_counter()
    s.advance(q, 200)
    s.synchronize()
    dt = time.perf_counter() - t0
    s.close()
    out["si_mach3_1486x1828"] = {"gpu_cell_updates_per_s": r.nx * r.ny * 200 / dt, "gpu_ms_per_step": dt / 200 * 1e3}
    # host round trip every step (fvm_set_state, one step, fvm_get_state) on the
    # headline grid: the rate when the caller keeps the state in host memory
    h = P.viscous_cylinder(nx=2048, ny=1024, diameter=64.0, reynolds=200.0, mach=0.3, center=(512.0, 512.0))
    h.dt_mode = P.DT_DEVICE_CFL
    s = fvm_amd.CFDSolver(0)
    s.initialize(h)
    q = h.copy()
    st = s.state()
    s.advance(q, 2)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        s.set_state(st)
        s.advance(q, 1)
        st = s.state()
    dt = time.perf_counter() - t0
    s.close()
    out["host_roundtrip_2048x1024"] = {"cell_updates_per_s": h.nx * h.ny * 10 / dt, "ms_per_step": dt / 10 * 1e3,
                                       "note": "state uploaded and read back (2 x 40 MB) around every step"}
    if with_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
        rate, sec = cpu_rate(p, 1000, threads)
        out["euler_256x128"].update({"cpu_cell_updates_per_s": rate, "cpu_seconds": sec, "cpu_threads": threads})
        cases = P.mixed_batch(512)[:32]
        cells = tot = 0.0
        for c in cases:
            r, sec = cpu_rate(c, 200, threads)
            cells += c.nx * c.ny * 200
            tot += sec
        out["mix_32x200_cpu"] = {"cpu_cell_updates_per_s": cells / tot, "cpu_seconds": tot, "cpu_threads": threads}
    return out


def batched_vs_plain(ncases, nsteps, graphs=True):
    """The mixed batch (grids 64x64..512x256, five shapes, Mach 0.3-4, half
    viscous, per-case CFL dt): all cases in the batched engine vs the same
    cases one at a time through the plain solver, nsteps each, same GPU."""
    import fvm_amd
    from fvm_amd import params as P
    from fvm_amd.batch import BatchEngine
    cases = P.mixed_batch(ncases)
    cells = sum(p.nx * p.ny for p in cases)
    # both sides capture their step graphs and take 2 warm-up steps before
    # their timed windows, so no capture is timed on either side
    b = BatchEngine(0, mem_budget_bytes=16 << 30, max_cases=ncases)
    b.use_graph(graphs)
    for p in cases:
        b.add(p)
    if graphs:
        b.prepare_graphs()
    b.advance(2)
    b.synchronize()
    t0 = time.perf_counter()
    b.advance(nsteps)
    b.synchronize()
    tb = time.perf_counter() - t0
    b.close()
    s = fvm_amd.CFDSolver(0)
    s.use_graph(graphs)
    tp = 0.0
    for p in cases:
        s.initialize(p)
        q = p.copy()
        if graphs:
            s.prepare_graphs(q)
        s.advance(q, 2)
        s.synchronize()
        t0 = time.perf_counter()
        s.advance(q, nsteps)
        s.synchronize()
        tp += time.perf_counter() - t0
    s.close()
    rb, rp = cells * nsteps / tb, cells * nsteps / tp
    return {"cases": ncases, "cells": cells, "steps": nsteps, "batched_cell_updates_per_s": rb,
            "plain_cell_updates_per_s": rp, "speedup": rb / rp}


def served(rate, seconds, step_scale, seed=5, time_limit=None):
    """Poisson request stream (grids 128x64..2048x1024 log-uniform, steps
    500-5000 x step_scale, 200 GB admission budget); latency = arrival to
    result read back."""
    from fvm_amd.serve import SolveServer, poisson_stream
    reqs = poisson_stream(rate, seconds, seed=seed, step_scale=step_scale)
    offered = sum(r.cells * r.steps for r in reqs) / seconds
    srv = SolveServer(0, mem_budget_bytes=200 << 30, max_active=1024, chunk_steps=25)
    st = srv.run(reqs, time_limit=time_limit if time_limit is not None else seconds * 3,
                 progress_every=30.0 if seconds > 60 else None)
    srv.close()
    st.pop("done")
    st.update({"rate_per_s": rate, "arrival_window_s": seconds, "step_scale": step_scale,
               "offered_cell_updates_per_s": offered})
    return st


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus {a.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import fvm_amd
    from fvm_amd import params as P

    p = P.viscous_cylinder(nx=a.nx, ny=a.ny, diameter=64.0, reynolds=200.0, mach=0.3,
                           center=(512.0, 512.0) if (a.nx, a.ny) == (2048, 1024) else None)
    p.dt_mode = P.DT_DEVICE_CFL
    s = fvm_amd.CFDSolver(0)
    s.use_graph(not a.direct)
    s.initialize(p)
    q = p.copy()
    # the 2-, 16- and 64-step graphs the timed steps replay are captured here, so
    # whatever --warmup is, the timed window holds no graph capture
    if not a.direct:
        s.prepare_graphs(q)
    # The GPU's clocks ramp up over the first few hundred steps after start-up
    # (a 64x7 stage kernel took 50.8 us after 20 steps and 44.9 us after 400),
    # so untimed steps run for --settle seconds before the --warmup steps;
    # the line reports how many.
    settle_steps = 0
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < a.settle:
        s.advance(q, 100)
        s.synchronize()
        settle_steps += 100
    s.advance(q, a.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.advance(q, a.steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    cells = p.nx * p.ny
    value = cells * a.steps / elapsed
    ms = elapsed / a.steps * 1e3

    # timing pass: per-kernel durations with HIP events on the solver's stream
    s.enable_timing(True)
    s.advance(q, a.timing_steps)
    s.synchronize()
    kt = s.kernel_timings()
    s.enable_timing(False)
    # The dominant kernel: the one that moves the most algorithmic bytes per
    # step (stage 2 of the split schedule: 61 of the step's 102 B per cell; the
    # two stage kernels also take the most time, within a few percent of each
    # other).  `per_kernel` gives every kernel's figures beside it.
    ran = [k for k in KERNEL_BYTES_PER_CELL if kt.get(k, (0, 0))[1] > 0]
    dominant = max(ran, key=lambda k: KERNEL_BYTES_PER_CELL[k] * kt[k][1])
    avg_ms = kt[dominant][0]
    alg_bytes = KERNEL_BYTES_PER_CELL[dominant] * cells
    achieved = alg_bytes / (avg_ms * 1e-3)
    prof = counters_from_profiles(a.nx, a.ny)
    traffic = prof.get("hbm_bytes_per_launch", {}).get(dominant)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": dominant,
                "kernel_avg_ms": avg_ms, "alg_bytes_per_launch": alg_bytes}
    per_kernel = {}
    for k in ran:
        kb = KERNEL_BYTES_PER_CELL[k] * cells
        per_kernel[k] = {"avg_ms": kt[k][0], "alg_bytes": kb, "hbm_frac": kb / (kt[k][0] * 1e-3) / HBM_PEAK}
        vk = prof.get("valu_insts_per_dispatch", {}).get(k)
        if vk:
            per_kernel[k]["valu_frac"] = vk * 64 / (kt[k][0] * 1e-3) / VALU_PEAK
    roofline["per_kernel"] = per_kernel
    # The roof that binds: vector-ALU issue.  SQ_INSTS_VALU (wave instructions
    # per launch, from the committed PMC summary of this build) x 64 lanes over
    # the same live launch duration, against 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz.
    valu = prof.get("valu_insts_per_dispatch", {}).get(dominant)
    if valu:
        lane_ops = valu * 64 / (avg_ms * 1e-3)
        roofline["binding"] = "valu"
        roofline["valu"] = {"achieved": lane_ops / 1e12, "peak": VALU_PEAK / 1e12, "unit": "lane-Tops/s",
                            "frac": lane_ops / VALU_PEAK, "insts_per_launch": valu}
        roofline["note"] = ("the stage kernels are vector-ALU issue bound: valu.frac is their VALU issue "
                            "fraction at the measured duration"
                            + (f"; HBM traffic is {traffic / alg_bytes:.2f}x the algorithmic bytes" if traffic else ""))

    t_sim, nsteps_dev = s.time()
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "clock_settle_steps": settle_steps,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"viscous Navier-Stokes cylinder {p.nx}x{p.ny}, Re 200, Mach 0.3, HLLC + van Leer, "
                               "SSP-RK2 + Strang viscous half steps, CFL 0.4 device dt",
                   "nx": p.nx, "ny": p.ny, "parallelism": "single GPU",
                   "launch": "direct" if a.direct else "hip graph"},
        "hbm_peak_fraction_at_102B": value * B_ALG_PER_CELL_UPDATE / HBM_PEAK,
        "roofline": roofline,
        "kernel_ms": {k: v[0] for k, v in kt.items()},
        "sim_time": t_sim,
    }
    s.close()
    if not a.no_batch:
        out["batched"] = batched_vs_plain(a.batch_cases, a.batch_steps, graphs=not a.direct)
    if not a.no_serve:
        out["served"] = served(a.serve_rate, a.serve_seconds, a.serve_step_scale, time_limit=a.serve_time_limit)
        # above capacity: the queue grows, latency is queueing; the run stops at 6x the window
        if a.overload_seconds > 0:
            out["served_overload"] = served(a.overload_rate, a.overload_seconds, a.serve_step_scale, seed=6,
                                            time_limit=6 * a.overload_seconds)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, a.cpu_steps)
    else:
        out["cpu_baseline"] = None
    if not a.no_other:
        out["other_configs"] = other_configs(not a.no_cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
