#!/usr/bin/env python3
"""Roe-RK throughput on one MI355X: cell-updates/s, FP32-mixed vs FP64.

Workload: 2D Euler Roe Kelvin-Helmholtz shear layer, 2048x2048 cells (ghosts
included) on [-0.5,0.5]^2, periodic, gamma 1.4, monotized-central limiter,
classic RK4, CFL 0.5; rho=2, vx=+.5 for |y|<.25 else rho=1, vx=-.5, p=2.5,
vy = .01 sin(4 pi x) + 1e-3 (2U-1) with U from a SplitMix64 counter hash.
One step = boundary + CFL reduction + 4 fused RK stages, dt on the device.

`value` is the FP32-mixed path's interior cell-updates per second over the K
timed steps with the state already in HBM.  The plain FP64 path is timed the
same way and reported beside it.  `roofline` covers the fused stage kernel:
algorithmic bytes per launch = interior cells x B / S with
B = N * 4 * (3S - 1) bytes per cell-update (N=4 states, S=4 stages, FP32
state), divided by the kernel's average duration measured with HIP events on
the simulation's stream.  `cpu_baseline` times the FP64 CPU restatement (the
oracle, OpenMP) on the host cores on a bounded number of steps of the same
grid.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))  # exact_riemann (closed-form Sod solution)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--kahan", type=int, default=0,
                    help="headline path with compensated FP32 state storage (1) or plain FP32 (0)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the Sod L1 accuracy leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the HBM "
                         "byte counters of a profiler run against a known transfer), then exit")
    return ap.parse_args()


def make_desc(n, precision, kahan=True):
    from hydrodynamicsgpu_amd import config

    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if precision == "fp32_mixed" and kahan else 0)
    return config.make_desc([n, n], xmin=[-0.5, -0.5], xmax=[0.5, 0.5], integratorName="RungeKutta4",
                            slopeLimiter="MonotizedCentral",
                            boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}] * 2, cfl=0.5,
                            defs={"idealGas_heatCapacityRatio": 1.4}, precision=precision, flags=flags)


def sync_all(sim):
    sim.sync()
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def time_gpu(precision, n, steps, warmup, q0, kahan=True):
    from hydrodynamicsgpu_amd import capi

    d = make_desc(n, precision, kahan)
    sim = capi.Sim(d)
    sim.set_state(q0)
    sim.run(warmup)
    sync_all(sim)
    # exactly `steps` timed steps, bracketed by synchronisation
    t0 = time.perf_counter()
    sim.run(steps)
    sync_all(sim)
    el_clean = time.perf_counter() - t0
    # then the same number of steps with HIP events around every fused stage
    # launch, for the per-kernel duration
    sim.set_profiling(True)
    t1 = time.perf_counter()
    sim.run(steps)
    sync_all(sim)
    el = time.perf_counter() - t1
    stage_ms, launches = sim.stage_timing()
    sim.set_profiling(False)
    health = sim.health()
    state = sim.get_state()
    t, nsteps, dt = sim.time()
    sim.close()
    return dict(elapsed=el_clean, elapsed_profiled=el, stage_ms=stage_ms, launches=launches, health=health,
                state=state, t=t, nsteps=nsteps)


def sod_l1(precision, kahan, n=1024):
    """L1 density error against the exact solution of the 1D Sod tube (Euler
    Roe, TVD-RK2, Superbee, free-flow, CFL 0.5, t = 0.2, 1024 cells on
    [-1, 1]) on the GPU path of the given precision."""
    from exact_riemann import sample
    from hydrodynamicsgpu_amd import capi, config, ics

    flags = config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND | (config.FLAG_KAHAN if kahan else 0)
    d = config.make_desc([n], xmin=[-1], xmax=[1], integratorName="RungeKutta2TVD", slopeLimiter="Superbee",
                         cfl=0.5, t_end=0.2, precision=precision, flags=flags)
    sim = capi.Sim(d)
    sim.set_state(ics.sod(d))
    while sim.time()[0] < 0.2 - 1e-12:
        sim.run(16)
    rho, _, _ = ics.primitives_euler(d, sim.get_state())
    x = ics.cell_centers(d)[0]
    er, _, _ = sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, x / 0.2)
    l1 = float(np.sum(np.abs(rho - er)[2:n - 2]) * (2.0 / n))
    sim.close()
    return l1


def cpu_baseline(n, seconds, q0):
    import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.lib().oracle_set_threads(threads)
    d = make_desc(n, "fp64")
    o = oracle.OracleSim(d)
    o.set_state(q0)
    o.update()  # first touch
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.update()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    cells = (n - 4) ** 2
    return dict(value=cells * steps / el, unit="cell-updates/s", cores=threads, kind="port",
                sample="%dx%d KH RK4 grid, %d steps after 1 untimed, FP64 oracle restatement, %d OpenMP threads"
                       % (n, n, steps, threads))


def copy_probe(mib):
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(one MI355X per box); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from hydrodynamicsgpu_amd import ics

    n = a.size
    N, S = 4, 4
    cells = (n - 4) ** 2
    q0 = ics.shear_kh(make_desc(n, "fp64"))

    mixed = time_gpu("fp32_mixed", n, a.steps, a.warmup, q0, kahan=bool(a.kahan))
    res = {}
    if not a.no_fp64:
        res["fp64"] = time_gpu("fp64", n, a.steps, a.warmup, q0)
        # the other FP32 storage form, for comparison
        res["other"] = time_gpu("fp32_mixed", n, a.steps, a.warmup, q0, kahan=not a.kahan)

    value = cells * a.steps / mixed["elapsed"]
    B = N * 4 * (3 * S - 1)  # algorithmic bytes per cell-update, FP32 state
    bytes_per_launch = cells * B / S
    achieved = bytes_per_launch / (mixed["stage_ms"] * 1e-3) / 1e9
    out = {
        "metric": "Roe-RK cell-updates/s, FP32-mixed vs FP64 at fixed L1 error; % HBM peak",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mixed["elapsed"] * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "euler2d_roe_kh", "grid": [n, n], "interior_cells": cells, "states": N,
                   "integrator": "RungeKutta4", "limiter": "MonotizedCentral", "boundary": "periodic",
                   "precision": "fp32_mixed+kahan" if a.kahan else "fp32_mixed",
                   "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "stage2d_kernel<RoePolicy<2,float,MonotizedCentral>,2,float,float,%s,mode>"
                               % ("true" if a.kahan else "false"),
                     "kernel_avg_ms": mixed["stage_ms"], "launches_timed": mixed["launches"],
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "health": mixed["health"],
    }
    if "fp64" in res:
        f = res["fp64"]
        v64 = cells * a.steps / f["elapsed"]
        out["fp64"] = {"value": v64, "ms_per_step": f["elapsed"] * 1e3 / a.steps, "stage_kernel_avg_ms": f["stage_ms"],
                       "achieved_GBs_fp64_bytes": cells * N * 8 * (3 * S - 1) / S / (f["stage_ms"] * 1e-3) / 1e9}
        out["mixed_over_fp64"] = value / v64
        # accuracy of the mixed path against the FP64 path after the same steps
        from hydrodynamicsgpu_amd import ics as _ics

        d = make_desc(n, "fp64")
        rho32 = mixed["state"].reshape(-1, N)[:, 0].reshape(n, n)[2:-2, 2:-2]
        rho64 = f["state"].reshape(-1, N)[:, 0].reshape(n, n)[2:-2, 2:-2]
        out["accuracy"] = {"l1_rho_mixed_minus_fp64": float(np.mean(np.abs(rho32 - rho64))),
                           "linf_rho_mixed_minus_fp64": float(np.max(np.abs(rho32 - rho64))),
                           "t_mixed": mixed["t"], "t_fp64": f["t"]}
        o = res["other"]
        out["fp32_" + ("plain" if a.kahan else "kahan")] = {
            "value": cells * a.steps / o["elapsed"], "ms_per_step": o["elapsed"] * 1e3 / a.steps,
            "stage_kernel_avg_ms": o["stage_ms"]}
    if not a.no_accuracy:
        # L1 error at fixed resolution against the exact Sod solution: the
        # mixed path must stay within 2x of the FP64 path
        l1 = {"fp64": sod_l1("fp64", False), "fp32_mixed": sod_l1("fp32_mixed", False),
              "fp32_mixed_kahan": sod_l1("fp32_mixed", True)}
        out["sod_l1"] = dict(l1, workload="euler1d_roe_sod_1024_rk2tvd_superbee_t0.2",
                             mixed_over_fp64=l1["fp32_mixed" + ("_kahan" if a.kahan else "")] / l1["fp64"])
    if not a.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(n, a.cpu_seconds, q0)
        except Exception as e:  # reported, never fatal for the GPU number
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
