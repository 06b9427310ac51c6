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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 512-grid leg")
    ap.add_argument("--no-long-accuracy", action="store_true",
                    help="skip the long-horizon error and drift legs (C2 t=.5, C3 t=.1, C4 t=.2 against 2x grids)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the HBM traffic measurement (rocprofv3 FETCH_SIZE / WRITE_SIZE child runs)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other configurations (Sod 1D, 3D HLLC blast, SRHD quadrants)")
    ap.add_argument("--only-config", choices=["C1", "C3", "C4"], default=None,
                    help="run only that configuration leg (Sod 1D, 3D HLLC blast, SRHD quadrants) and print "
                         "its JSON (for profiling one configuration)")
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
    phases = sim.phase_timing()
    sim.set_profiling(False)
    health = sim.health()
    state = sim.get_state()
    t, nsteps, dt = sim.time()
    # host-inclusive rates: upload of the FP64 AoS host array, K (or 1)
    # steps, readback, i.e. what a caller that keeps its state on the host pays
    host = {}
    for k in (steps, 1):
        t2 = time.perf_counter()
        sim.set_state(q0)
        sim.run(k)
        sim.get_state()
        host[k] = time.perf_counter() - t2
    sim.close()
    return dict(elapsed=el_clean, elapsed_profiled=el, stage_ms=stage_ms, launches=launches, health=health,
                state=state, t=t, nsteps=nsteps, host=host, phases=phases)


def sod_l1(precision, kahan, n=1024):
    """L1 density error against the exact solution of the 1D Sod tube (Euler
    Roe, TVD-RK2, Superbee, free-flow, CFL 0.5, t = 0.2, 1024 cells on
    [-1, 1]) on the GPU path of the given precision."""
    from hydrodynamicsgpu_amd.riemann import sample
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


def batch_leg(steps, warmup):
    """The uneven 512-grid mix (1D lines of 256..65536 cells and 2D squares of
    64^2..512^2, Riemann problems, Euler Roe, Superbee, RK2, FP32-mixed) in
    one batch, after the autotuner picked its decomposition, against the
    single 2048^2 grid with the same scheme (the large-grid throughput)."""
    from hydrodynamicsgpu_amd import capi, config, ics

    kw = dict(integratorName="RungeKutta2", slopeLimiter="Superbee", precision="fp32_mixed")
    mix = ics.riemann_batch(512, **kw)
    b = capi.Batch([d for d, _ in mix])
    for m, (_, q) in enumerate(mix):
        b.set_state(m, q)
    rows, strips, _ = b.autotune()
    b.run(warmup)
    b.sync()
    t0 = time.perf_counter()
    b.run(steps)
    b.sync()
    el = time.perf_counter() - t0
    cells = b.interior_cells()
    health = max(b.health(m) for m in range(len(mix)))
    b.close()
    # the large grid, same scheme
    n = 2048
    d = config.make_desc([n, n], xmin=[-0.5, -0.5], xmax=[0.5, 0.5],
                         boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}] * 2, cfl=0.5, **kw)
    sim = capi.Sim(d)
    sim.set_state(ics.shear_kh(d))
    sim.run(warmup)
    sim.sync()
    t1 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    el1 = time.perf_counter() - t1
    sim.close()
    big = (n - 4) ** 2 * steps / el1
    v = cells * steps / el
    return {"workload": "512 uneven Euler Roe grids (1D 256..65536, 2D 64^2..512^2), Superbee, RK2, fp32_mixed",
            "members": len(mix), "interior_cells": cells, "value": v, "unit": "cell-updates/s",
            "ms_per_step": el * 1e3 / steps, "decomposition": {"rows": rows, "strips_1d": strips},
            "large_grid_value": big, "large_grid": "2048^2 KH, same scheme", "fraction_of_large_grid": v / big,
            "health": health}


def time_desc(d, q0, steps, warmup):
    """Timed run of one configuration on the GPU: `steps` updates after
    `warmup`, bracketed by synchronisation, then the same with HIP events on
    the stage launches."""
    from hydrodynamicsgpu_amd import capi

    sim = capi.Sim(d)
    sim.set_state(q0)
    sim.run(warmup)
    sync_all(sim)
    t0 = time.perf_counter()
    sim.run(steps)
    sync_all(sim)
    el = time.perf_counter() - t0
    sim.set_profiling(True)
    sim.run(max(2, steps // 4))
    stage_ms, launches = sim.stage_timing()
    sim.set_profiling(False)
    out = dict(elapsed=el, stage_ms=stage_ms, launches=launches, health=sim.health(), state=sim.get_state())
    sim.close()
    return out


def host_cpu():
    """The host the CPU baseline runs on: logical CPUs (nproc), the CPUs this
    process may run on, OMP_NUM_THREADS, and the CPU model and socket count
    from /proc/cpuinfo (what lscpu reports)."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        info["affinity_cpus"] = None
    info["omp_num_threads"] = os.environ.get("OMP_NUM_THREADS")
    model, sockets = None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    info["model"] = model
    info["sockets"] = len(sockets) or None
    return info


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may use, or
    OMP_NUM_THREADS when the environment sets it (a GPU job's share of a
    shared host)."""
    info = host_cpu()
    share = info["affinity_cpus"] or info["nproc"] or 1
    env = info["omp_num_threads"]
    if env and env.isdigit() and int(env) > 0:
        return min(int(env), share)
    return share


def cpu_steps(d, q0, seconds, max_steps, threads=None):
    """The FP64 oracle on the host cores: cell-updates/s over a bounded number of steps."""
    import oracle

    threads = threads or cpu_threads()
    oracle.lib().oracle_set_threads(threads)
    o = oracle.OracleSim(d)
    o.set_state(q0)
    o.update()
    n = 0
    t0 = time.perf_counter()
    while True:
        o.update()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= max_steps:
            break
    return n, el, threads


def _pmc(counter, args, outdir):
    """One rocprofv3 counter pass over a child run of this script; returns
    {kernel name: [value per dispatch]} (KiB for FETCH_SIZE / WRITE_SIZE)."""
    import csv
    import glob
    import shutil
    import subprocess

    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", outdir, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + args
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=120)
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if r.returncode != 0 or not files:
        return None
    out = {}
    for row in csv.DictReader(open(files[0])):
        if row["Counter_Name"] == counter:
            out.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return out


def traffic_leg(kahan):
    """HBM bytes per stage launch: FETCH_SIZE and WRITE_SIZE of the stage
    kernels in child runs of the headline workload, each scaled by what the
    same counter reads for a plain 256 MiB device copy (known bytes)."""
    import tempfile

    work = tempfile.mkdtemp(prefix="hydro_pmc_", dir="/tmp")
    base = ["--steps", "3", "--warmup", "1", "--no-cpu", "--no-fp64", "--no-accuracy", "--no-batch", "--no-configs",
            "--no-traffic", "--no-long-accuracy", "--kahan", str(int(kahan))]
    probe = ["--copy-probe", "256"]
    f = _pmc("FETCH_SIZE", base, os.path.join(work, "f"))
    w = _pmc("WRITE_SIZE", base, os.path.join(work, "w"))
    cf = _pmc("FETCH_SIZE", probe, os.path.join(work, "cf"))
    cw = _pmc("WRITE_SIZE", probe, os.path.join(work, "cw"))
    if not (f and w and cf and cw):
        return None
    known = 256.0 * (1 << 20)

    def probe_kib(d):
        vals = [v for k, vs in d.items() if "copyBuffer" in k or "copy" in k.lower() for v in vs]
        return max(vals) if vals else None

    cfk, cwk = probe_kib(cf), probe_kib(cw)
    if not cfk or not cwk:
        return None
    ff = known / (cfk * 1024.0)
    wf = known / (cwk * 1024.0)
    fs = [v for k, vs in f.items() if "stage2d_kernel" in k for v in vs]
    ws = [v for k, vs in w.items() if "stage2d_kernel" in k for v in vs]
    if not fs or not ws:
        return None
    fb = sum(fs) / len(fs) * 1024.0 * ff
    wb = sum(ws) / len(ws) * 1024.0 * wf
    return {"bytes_per_launch": fb + wb, "fetch_bytes": fb, "write_bytes": wb, "fetch_scale": ff, "write_scale": wf,
            "launches": len(fs)}


def config_legs(steps, warmup, with_cpu, cpu_seconds, only=None):
    """The other configurations: Sod 1D (FP64, to t = 0.2, L1 vs the exact
    solution), the 3D HLLC blast (256^3, TVD-RK3, mixed with Kahan vs FP64)
    and the SRHD quadrants (1024^2, RK2, FP32 fluxes + FP64 con2prim vs FP64).
    B per cell-update: N b (3S - 1), SRHD + 4 N b for the primitives (stored
    in the state's precision b)."""
    from hydrodynamicsgpu_amd import capi, config, ics
    from hydrodynamicsgpu_amd.riemann import sample

    res = {}
    if only not in (None, "C1"):
        return _config_pairs(res, steps, warmup, with_cpu, cpu_seconds, only)
    # C1: Sod, 1024 cells incl. ghosts on [-1, 1], TVD-RK2, Superbee, CFL 0.5, t = 0.2
    n = 1024
    kw = dict(xmin=[-1], xmax=[1], integratorName="RungeKutta2TVD", slopeLimiter="Superbee", cfl=0.5, t_end=0.2)
    x = ics.cell_centers(config.make_desc([n], **kw))[0]
    er, _, _ = sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, x / 0.2)

    def sod_to_t(precision, flags):
        d = config.make_desc([n], precision=precision, flags=config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND | flags,
                             **kw)
        q = ics.sod(d)
        sim = capi.Sim(d)
        sim.set_state(q)
        sim.run(2)
        sim.set_state(q)
        tot0 = sim.totals()
        sync_all(sim)
        t0 = time.perf_counter()
        while sim.time()[0] < 0.2 - 1e-12:
            sim.run(64)
        el = time.perf_counter() - t0
        _, nsteps, _ = sim.time()
        rho, _, _ = ics.primitives_euler(d, sim.get_state())
        acc = {"l1": float(np.sum(np.abs(rho - er)[2:n - 2]) * (2.0 / n)),
               "linf": float(np.max(np.abs(rho - er)[2:n - 2])),
               "drift": _drift({"totals0": tot0, "totals1": sim.totals()})}
        sim.close()
        return d, q, nsteps, el, acc

    d, q, nsteps, el, acc64 = sod_to_t("fp64", 0)
    _, _, nsteps32, el32, acc32 = sod_to_t("fp32_mixed", config.FLAG_KAHAN)
    l1, l1_32 = acc64["l1"], acc32["l1"]
    c1 = {"workload": "euler1d_roe_sod_1024_rk2tvd_superbee_t0.2", "precision": "fp64", "steps": nsteps,
          "gpu_value": (n - 4) * nsteps / el, "unit": "cell-updates/s", "gpu_seconds_to_t": el, "l1_rho_vs_exact": l1,
          "gpu_mixed_value": (n - 4) * nsteps32 / el32, "gpu_mixed_seconds_to_t": el32, "mixed_steps": nsteps32,
          "l1_rho_vs_exact_mixed": l1_32,
          "accuracy": {"t_end": 0.2, "mixed_form": "fp32_mixed+kahan", "reference": "exact Riemann solution",
                       "l1": {"fp64": l1, "mixed": l1_32}, "l1_ratio": l1_32 / l1,
                       "linf": {"fp64": acc64["linf"], "mixed": acc32["linf"]},
                       "drift": {"fp64": acc64["drift"], "mixed": acc32["drift"]}},
          "note": "latency-bound: 1020 cells per step, every update in one single-workgroup launch per 64 steps, the host "
                  "reads t between launches; mixed = FP32 compensated storage"}
    if with_cpu:
        import oracle

        # 1020 cells per step: a thread team's fork and join per kernel can
        # cost more than the work, so both one thread and the host share are
        # timed and the faster is the baseline
        runs = {}
        for th in sorted({1, cpu_threads()}):
            oracle.lib().oracle_set_threads(th)
            o = oracle.OracleSim(d)
            o.set_state(q)
            ocount = 0
            t0 = time.perf_counter()
            while o.time < 0.2 - 1e-14:
                o.update()
                ocount += 1
            runs[th] = (time.perf_counter() - t0, ocount)
            o.close()
        best = min(runs, key=lambda k: runs[k][0])
        c1["cpu_seconds_to_t"] = runs[best][0]
        c1["cpu_value"] = (n - 4) * runs[best][1] / runs[best][0]
        c1["cpu_cores"] = best
        c1["cpu_seconds_by_threads"] = {str(k): v[0] for k, v in runs.items()}
    res["C1"] = c1
    if only == "C1":
        return res
    return _config_pairs(res, steps, warmup, with_cpu, cpu_seconds, None)


def _config_pairs(res, steps, warmup, with_cpu, cpu_seconds, only):
    """C3 and C4, each timed in FP64 and FP32-mixed (`only`: one of them)."""
    from hydrodynamicsgpu_amd import config, ics

    def pair(name, workload, make, ic, N, S, b_extra=0, mixed_flags=0, stage_extra=0):
        # b_extra / stage_extra: values (not bytes) per cell-update / per
        # stage launch and cell beyond the state, at the state's width b
        if only not in (None, name):
            return
        d64 = make("fp64", 0)
        q0 = ic(d64)
        cells = int(np.prod([sz - 4 for sz in list(d64.size)[:d64.dim]]))
        g64 = time_desc(d64, q0, steps, warmup)
        g32 = time_desc(make("fp32_mixed", mixed_flags), q0, steps, warmup)
        v64 = cells * steps / g64["elapsed"]
        v32 = cells * steps / g32["elapsed"]
        B32 = (N * (3 * S - 1) + b_extra) * 4
        B64 = (N * (3 * S - 1) + b_extra) * 8
        # a stage launch moves its own algorithmic bytes: N b (3S - 1) / S
        # of state, plus what the stage itself reads besides (SRHD: the
        # primitives of the start-of-step state, N b); the rest of b_extra
        # belongs to the kernels between the stages and counts at step level
        a32 = cells * (N * (3 * S - 1) / S + stage_extra) * 4 / (g32["stage_ms"] * 1e-3) / 1e9
        a64 = cells * (N * (3 * S - 1) / S + stage_extra) * 8 / (g64["stage_ms"] * 1e-3) / 1e9
        ms32 = g32["elapsed"] * 1e3 / steps
        ms64 = g64["elapsed"] * 1e3 / steps
        diff = np.abs(g32["state"] - g64["state"]).reshape(-1, N)[:, 0]
        out = {"workload": workload, "interior_cells": cells, "gpu_fp64_value": v64, "gpu_mixed_value": v32,
               "unit": "cell-updates/s", "mixed_over_fp64": v32 / v64,
               "ms_per_step": {"fp64": ms64, "mixed": ms32},
               "stage_kernel_avg_ms": {"fp64": g64["stage_ms"], "mixed": g32["stage_ms"]},
               "roofline_frac_mixed": a32 / HBM_PEAK_GBS, "roofline_frac_fp64": a64 / HBM_PEAK_GBS,
               "roofline_frac_step_mixed": cells * B32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "roofline_frac_step_fp64": cells * B64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "B_mixed": B32, "B_fp64": B64, "l1_rho_mixed_minus_fp64_after_steps": float(np.mean(diff)),
               "health": {"fp64": g64["health"], "mixed": g32["health"]}}
        if with_cpu:
            ns, el, th = cpu_steps(d64, q0, cpu_seconds, 3)
            out["cpu_value"] = cells * ns / el
            out["cpu_cores"] = th
        res[name] = out

    def c3(prec, flags):
        return config.make_desc([256, 256, 256], xmin=[-1] * 3, xmax=[1] * 3, solverName="EulerHLLC",
                                integratorName="RungeKutta3TVD", slopeLimiter="MinMod",
                                boundaryMethods=[{"min": "MIRROR", "max": "MIRROR"}] * 3, cfl=0.4, precision=prec,
                                flags=config.FLAG_REFERENCE | flags)

    pair("C3", "euler3d_hllc_blast_256^3_rk3tvd", c3, ics.blast, 5, 3, mixed_flags=config.FLAG_KAHAN)

    def c4(prec, flags):
        return config.make_desc([1024, 1024], xmin=[0, 0], xmax=[1, 1], solverName="SRHDRoe",
                                integratorName="RungeKutta2", defs={"idealGas_heatCapacityRatio": 5.0 / 3.0},
                                precision=prec, flags=config.FLAG_REFERENCE | flags)

    pair("C4", "srhd2d_roe_quadrants_1024^2_rk2", c4, ics.srhd_quadrants, 4, 2, b_extra=4 * 4, stage_extra=4)
    if "C4" in res:
        from hydrodynamicsgpu_amd import capi

        sim = capi.Sim(c4("fp64", 0))
        sim.set_state(ics.srhd_quadrants(sim.desc))
        P = sim.get_primitives().reshape(-1, 4)
        sim.close()
        res["C4"]["max_lorentz_factor_initial"] = float(1.0 / np.sqrt(1.0 - (P[:, 1] ** 2 + P[:, 2] ** 2).max()))
    return res


def _progress(msg):
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def _run_to(desc, q0, t_end, chunk=128):
    """One grid from q0 to t_end (the last step clipped to it): the final
    state, its FP64 conserved totals at the start and the end, and the clock."""
    from hydrodynamicsgpu_amd import capi

    sim = capi.Sim(desc)
    sim.set_state(q0)
    tot0 = sim.totals()
    while sim.time()[0] < t_end * (1 - 1e-13):
        sim.run(chunk)
    t, steps, _ = sim.time()
    out = dict(state=sim.get_state(), totals0=tot0, totals1=sim.totals(), t=t, steps=steps, health=sim.health())
    sim.close()
    return out


def _interior_rho(desc, aos):
    N = desc.dim + 2
    shape = tuple(desc.size[a] for a in reversed(range(desc.dim)))
    rho = np.asarray(aos).reshape(-1, N)[:, 0].reshape(shape)
    return rho[(slice(2, -2),) * desc.dim]


def _restrict(fine, dim, factor=2):
    """Block average of a fine interior onto the coarse interior."""
    shp = []
    for n in fine.shape:
        shp += [n // factor, factor]
    return fine.reshape(shp).mean(axis=tuple(range(1, 2 * dim, 2)))


def _drift(r):
    t0, t1 = np.asarray(r["totals0"]), np.asarray(r["totals1"])
    rel = [float((b - a) / abs(a)) if abs(a) > 1e-12 * abs(t0[0]) else None for a, b in zip(t0, t1)]
    return {"relative": rel, "absolute": [float(b - a) for a, b in zip(t0, t1)]}


def accuracy_legs(only=None):
    """Error and conservation at long horizons, per configuration: the FP32
    mixed and the FP64 GPU runs of the scheme against an FP64 run on a grid
    twice as fine over the same interior region (the same initial
    condition: the refined grid's noise is its parent cell's), restricted to
    the base grid by block averaging; density L1 (integral of |e| over the
    interior) and L-infinity; and the drift of the FP64 conserved totals
    (interior sums) from t = 0 to t_end.  The FP64 GPU path stands for the
    FP64 oracle here (it matches it to 1e-10 in the parity tests); running
    the CPU oracle this far would take hours."""
    from hydrodynamicsgpu_amd import config, ics

    def kh(prec, flags, t_end):
        return config.make_desc([2048, 2048], xmin=[-0.5, -0.5], xmax=[0.5, 0.5], integratorName="RungeKutta4",
                                slopeLimiter="MonotizedCentral",
                                boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}] * 2, cfl=0.5,
                                precision=prec, flags=config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND | flags,
                                t_end=t_end)

    def blast(prec, flags, t_end):
        return config.make_desc([256, 256, 256], xmin=[-1] * 3, xmax=[1] * 3, solverName="EulerHLLC",
                                integratorName="RungeKutta3TVD", slopeLimiter="MinMod",
                                boundaryMethods=[{"min": "MIRROR", "max": "MIRROR"}] * 3, cfl=0.4, precision=prec,
                                flags=config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND | flags, t_end=t_end)

    def quads(prec, flags, t_end):
        return config.make_desc([1024, 1024], xmin=[0, 0], xmax=[1, 1], solverName="SRHDRoe",
                                integratorName="RungeKutta2", defs={"idealGas_heatCapacityRatio": 5.0 / 3.0},
                                precision=prec, flags=config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND | flags,
                                t_end=t_end)

    legs = [
        ("C2", kh, 0.5, lambda d, cells=None: ics.shear_kh(d, cells=cells), 0, "fp32_mixed"),
        ("C3", blast, 0.1, lambda d, cells=None: ics.blast(d, cells=cells), config.FLAG_KAHAN,
         "fp32_mixed+kahan"),
        ("C4", quads, 0.2, lambda d, cells=None: ics.srhd_quadrants(d), 0, "fp32_mixed"),
    ]
    res = {}
    for name, make, t_end, ic, mflags, mname in legs:
        if only not in (None, name):
            continue
        t0 = time.perf_counter()
        d64 = make("fp64", 0, t_end)
        q = ic(d64)
        r64 = _run_to(d64, q, t_end)
        r32 = _run_to(make("fp32_mixed", mflags, t_end), q, t_end)
        del q
        df = ics.refined(d64)
        qf = ic(df, cells=ics.parent_cells(d64, df))
        rf = _run_to(df, qf, t_end)
        del qf
        ref = _restrict(_interior_rho(df, rf["state"]), d64.dim)
        dV = float(np.prod([(d64.xmax[a] - d64.xmin[a]) / d64.size[a] for a in range(d64.dim)]))
        e64 = _interior_rho(d64, r64["state"]) - ref
        e32 = _interior_rho(d64, r32["state"]) - ref
        l1_64, l1_32 = float(np.sum(np.abs(e64)) * dV), float(np.sum(np.abs(e32)) * dV)
        dm = _interior_rho(d64, r32["state"]) - _interior_rho(d64, r64["state"])
        res[name] = {
            "t_end": t_end, "mixed_form": mname, "steps": {"fp64": r64["steps"], "mixed": r32["steps"],
                                                          "reference_2x": rf["steps"]},
            "reference": "FP64 GPU run at 2x resolution (%s cells incl. ghosts), block-averaged"
                         % "x".join(str(df.size[a]) for a in range(df.dim)),
            "l1": {"fp64": l1_64, "mixed": l1_32}, "l1_ratio": l1_32 / l1_64,
            "linf": {"fp64": float(np.max(np.abs(e64))), "mixed": float(np.max(np.abs(e32)))},
            "mixed_minus_fp64": {"l1": float(np.sum(np.abs(dm)) * dV), "linf": float(np.max(np.abs(dm)))},
            "drift": {"fp64": _drift(r64), "mixed": _drift(r32), "reference_2x": _drift(rf)},
            "health": {"fp64": r64["health"], "mixed": r32["health"], "reference_2x": rf["health"]},
        }
        del rf, r64, r32
        _progress("accuracy %s to t=%g: %.1f s" % (name, t_end, time.perf_counter() - t0))
    return res


def cpu_baseline(n, seconds, q0):
    import oracle

    threads = cpu_threads()
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
                       % (n, n, steps, threads), host=host_cpu())


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
    if a.only_config:
        print(json.dumps(config_legs(a.steps, a.warmup, not a.no_cpu, min(a.cpu_seconds, 10.0), a.only_config)))
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
        # device time per update by phase (HIP events while profiling; the
        # profiled run launches kernel by kernel, so its sum exceeds
        # ms_per_step by the launch gaps the graph replay hides)
        "phases_us_per_update": {k: round(v[0] * 1e3 / a.steps, 2)
                                 for k, v in mixed["phases"].items() if v[1]},
        "host_inclusive": {
            "note": "not the headline: set_state from a host FP64 AoS array + steps + get_state, wall clock",
            "value_k_steps": cells * a.steps / mixed["host"][a.steps],
            "value_1_step": cells / mixed["host"][1],
            "ms_transfer_overhead": (mixed["host"][a.steps] - mixed["elapsed"]) * 1e3},
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
    if not a.no_traffic:
        try:
            t = traffic_leg(a.kahan)
        except Exception as e:  # reported, never fatal for the GPU number
            t = None
            out["roofline"]["traffic_error"] = str(e)
        if t:
            out["roofline"]["traffic"] = t["bytes_per_launch"]
            out["roofline"]["traffic_detail"] = t
    _progress("headline legs done")
    if not a.no_batch:
        out["batch"] = batch_leg(a.steps, a.warmup)
        _progress("batch leg done")
    if not a.no_configs:
        out["configs"] = config_legs(a.steps, a.warmup, not a.no_cpu, min(a.cpu_seconds, 10.0))
        _progress("configuration legs done")
        c2 = {"workload": "euler2d_roe_kh_2048^2_rk4_mc", "interior_cells": cells, "gpu_mixed_value": value,
              "unit": "cell-updates/s", "ms_per_step": {"mixed": out["ms_per_step"]},
              "roofline_frac_mixed": out["roofline"]["frac"],
              "roofline_frac_step_mixed": cells * B / (out["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
              "B_mixed": B, "B_fp64": N * 8 * (3 * S - 1)}
        if "fp64" in out:
            c2.update(gpu_fp64_value=out["fp64"]["value"], mixed_over_fp64=out["mixed_over_fp64"])
            c2["ms_per_step"]["fp64"] = out["fp64"]["ms_per_step"]
            c2["roofline_frac_step_fp64"] = (cells * c2["B_fp64"] / (out["fp64"]["ms_per_step"] * 1e-3) / 1e9
                                             / HBM_PEAK_GBS)
        out["configs"] = dict({"C2": c2}, **out["configs"])
    if not a.no_long_accuracy:
        acc = accuracy_legs()
        out.setdefault("configs", {})
        for k, v in acc.items():
            out["configs"].setdefault(k, {})["accuracy"] = v
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
