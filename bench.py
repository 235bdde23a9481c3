"""Throughput benchmark of the FlowLB core on one MI355X.

Default workload: the 8192 x 2048 free-surface culvert (dam break behind an
embankment with a 240-cell barrel, gravity 2e-5, tau = 0.51), FP32 population
storage.  A step is one full free-surface lattice step (passes 1-4 around
stream-collide) over the whole grid, inputs resident in HBM.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--storage fp32]

Prints one JSON line: MLUPS over all cells, the roofline of the dominant
kernel (stream-collide) from HIP-event timing of its launches, the FP16/BF16
storage variants and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS & % of HBM roofline, FP32 vs FP16-store; adjoint s/gradient"
HBM_PEAK_GBS = 8000.0
STORAGE_BYTES = {"fp32": 4, "fp16s": 2, "bf16s": 2}


def active_counts(solver):
    """(active, inactive) cells from the device flag histogram: WALL and any
    GAS-flagged cell are inactive."""
    h = solver.flag_histogram()
    vals = np.arange(256)
    inactive = h[(vals & 0x0A) != 0].sum()
    return int(h.sum() - inactive), int(inactive)


def algorithmic_bytes(active, inactive, s, free_surface):
    """Per-launch bytes of stream-collide and per-step bytes of the whole step.
    Stream-collide: 18 s (nine loads, nine stores) + 1 flag byte per active cell,
    + 13 in (conservative) free-surface runs, as a body-of-fluid cell -- whose
    update stream-collide does -- moves them: the excess shares pulled from
    the neighbours (each share is read by 8 neighbours, 4 bytes per cell
    once), the mass store, its new excess share and flag (a surface cell
    stores its density instead of the share and flag: 1 byte less); the
    interface cells' extra velocity store and reconstructed populations
    (< 0.2 % of the cells) are left out; 1 flag byte per inactive cell.  With
    16-bit storage a free-surface cell also reads and writes its 2-byte
    rounding residual (the mass the population stores rounded away).
    Free-surface step: 18 s + 38 per active cell (+ 4 with 16-bit storage)."""
    resid = 4 if free_surface and s < 4 else 0
    sc = active * (18 * s + 1 + (13 if free_surface else 0) + resid) + inactive
    step = active * (18 * s + (38 if free_surface else 1) + resid) + inactive
    return sc, step


def _oracle_rate(case, fp32, threads, seconds):
    """MLUPS of the oracle on the full grid of the workload: steps after one
    warm-up step until `seconds` have passed (at most 1000)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    flags, rho0, u0 = case.build()
    o = pyoracle.Oracle(flags, case.nx, case.ny, case.tau, rho0=rho0.astype(np.float64),
                        u0=u0.astype(np.float64), use_force=case.force != (0.0, 0.0),
                        fx=case.force[0], fy=case.force[1], free_surface=case.free_surface,
                        mass_mode=1, threads=threads, fp32=fp32)
    o.step(1)
    n, t0 = 0, time.perf_counter()
    while True:
        o.step(1)
        n += 1
        dt = time.perf_counter() - t0
        if dt > seconds or n >= 1000:
            break
    o.close()
    return case.N * n / dt / 1e6, n, dt


def cpu_baseline(case, seconds=10.0):
    """The oracle timed on this host: FP32 with OpenMP on all cores of the
    host's share (the headline CPU figure), FP32 on one core (for the parallel
    efficiency) and FP64 on one core, each on a bounded sample of the
    workload's full grid; plus the 512 x 128 single-phase culvert run in full
    (10^4 steps, FP64, one core)."""
    from flowlb_amd import cases
    cores = min(16, os.cpu_count() or 1)
    par, npar, dpar = _oracle_rate(case, True, cores, seconds)
    one, n1, d1 = _oracle_rate(case, True, 1, seconds)
    f64, n64, d64 = _oracle_rate(case, False, 1, seconds)
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    c1 = cases.SINGLE_PHASE
    flags, rho0, u0 = c1.build()
    o = pyoracle.Oracle(flags, c1.nx, c1.ny, c1.tau, rho0=rho0.astype(np.float64),
                        u0=u0.astype(np.float64), threads=1)
    t0 = time.perf_counter()
    o.step(c1.steps)
    c1_s = time.perf_counter() - t0
    o.close()
    return dict(
        value=round(par, 3), unit="MLUPS", cores=cores, kind="port",
        sample=f"FP32 oracle with OpenMP on {cores} threads, {npar} steps of the full "
               f"{case.nx}x{case.ny} grid after 1 warm-up step ({dpar:.1f} s)",
        fp32_one_core={"value": round(one, 3), "steps": n1, "seconds": round(d1, 2)},
        parallel_efficiency=round(par / (one * cores), 3),
        fp64_one_core={"value": round(f64, 3), "steps": n64, "seconds": round(d64, 2)},
        single_phase_512x128_full_run={"steps": c1.steps, "seconds": round(c1_s, 2),
                                       "MLUPS": round(c1.N * c1.steps / c1_s / 1e6, 3),
                                       "precision": "fp64", "cores": 1})


def measure(case, storage, steps, warmup, profile_steps, t0=0, **opt):
    """Advances the workload to lattice time t0 and then `warmup` steps, times
    `steps` steps (HIP events on the solver stream), then `profile_steps` with
    every kernel bracketed by events (the per-kernel means)."""
    s = case.solver(storage, **opt)
    try:
        if t0:
            s.step(t0)
        s.step(warmup)
        s.sync()
        a0, i0 = active_counts(s)
        t_start = s.time
        w0 = time.perf_counter()
        ms = s.step_timed(steps)
        wall = time.perf_counter() - w0
        a1, i1 = active_counts(s)
        active, inactive = (a0 + a1) / 2, (i0 + i1) / 2
        tot_k, per_k = s.step_timed(profile_steps, per_kernel=True)
        mass_ok = np.isfinite(s.mass())
    finally:
        s.close()
    sbytes = STORAGE_BYTES[storage]
    sc_b, step_b = algorithmic_bytes(active, inactive, sbytes, case.free_surface)
    sc_ms = per_k[0] / profile_steps
    return dict(ms=ms, wall=wall, active=active, inactive=inactive, sc_bytes=sc_b,
                step_bytes=step_b, sc_ms=sc_ms, per_kernel_ms=[v / profile_steps for v in per_k],
                finite=bool(mass_ok), t_start=int(t_start))


def hbm_reference():
    """Rates a plain device copy and an in-place update (x += 1) of a 1 GiB FP32
    array reach on this GPU (torch's kernels, 10 launches each, events on
    torch's stream): what this device's HBM delivers in practice, next to the
    published 8 TB/s peak."""
    import torch
    n = (1 << 30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    out = {}
    for key, fn in (("copy_GBs", lambda: y.copy_(x)), ("inplace_update_GBs", lambda: x.add_(1.0))):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        out[key] = round(2 * 4 * n / (ms * 1e-3) / 1e9, 1)
    del x, y
    torch.cuda.empty_cache()
    return out


def hbm_reference_child():
    """hbm_reference() in a child process: torch brings its own HIP runtime,
    which does not start in a process where the solver library's has."""
    import subprocess
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--workload", "hbm_reference"],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        return {"error": r.stderr.strip().splitlines()[-1] if r.stderr.strip() else "failed"}
    return json.loads(r.stdout.strip().splitlines()[-1])


def brinkman_window(case, seed=None):
    x0, x1, y0, y1 = case.extra["window"]
    rng = np.random.default_rng(case.seed if seed is None else seed)
    alpha = np.zeros((case.ny, case.nx), np.float32)
    alpha[y0:y1 + 1, x0:x1 + 1] = rng.uniform(0, 1, (y1 - y0 + 1, x1 - x0 + 1))
    return alpha.ravel()


def bench_adjoint(args):
    """Gradient of the time-averaged outlet discharge w.r.t. the 64x64 Brinkman
    window and the inlet velocity on the 1024x256 culvert: GPU time per
    gradient, cost relative to the forward run, and accuracy against the FP64
    oracle adjoint and central finite differences of the FP64 objective."""
    from flowlb_amd import cases, core
    case = cases.ADJOINT
    steps = case.steps
    t0, t1 = case.extra["t_avg"]
    col = case.nx - 2
    ckpt = case.extra["ckpt"]
    flags, rho0, u0 = case.build()
    alpha = brinkman_window(case)
    with case.solver() as s:
        s.set_brinkman(alpha)
        s.objective(200, col, 0, 199)            # warm-up
        plain_ms = s.step_timed(steps)
        s.close()
    with case.solver() as s:
        s.set_brinkman(alpha)
        t = time.perf_counter()
        J_fwd = s.objective(steps, col, t0, t1)
        obj_s = time.perf_counter() - t
        s.adjoint(200, col, 100, 199, ckpt)      # warm-up
        t = time.perf_counter()
        J, g, duin, phases = s.adjoint(steps, col, t0, t1, ckpt)
        adj_s = time.perf_counter() - t
    # The same gradient from an FP64 handle: forward, trace and adjoint in FP64.
    with case.solver("fp64") as s:
        s.set_brinkman(alpha)
        s.adjoint(200, col, 100, 199, ckpt)      # warm-up
        t = time.perf_counter()
        J64, g64, duin64, phases64 = s.adjoint(steps, col, t0, t1, ckpt)
        adj64_s = time.perf_counter() - t
    out = {"metric": METRIC, "value": round(adj_s, 4), "unit": "s/gradient", "n_gpus": 1,
           "steps": steps, "warmup": 1, "ms_per_step": round(adj_s * 1e3 / steps, 5),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic",
           "config": {"workload": case.name, "nx": case.nx, "ny": case.ny, "window": "64x64",
                      "checkpoint_every": ckpt, "average_steps": [t0, t1]},
           "J": J, "forward_ms": round(plain_ms, 3),
           "adjoint_over_forward": round(adj_s * 1e3 / plain_ms, 3),
           "phases_ms": [round(v, 3) for v in phases],
           "fp64_handle": {"s_per_gradient": round(adj64_s, 4),
                           "phases_ms": [round(v, 3) for v in phases64], "J": J64}}
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle
        a64 = alpha.astype(np.float64)
        mk = lambda a, uin=None: pyoracle.Oracle(  # noqa: E731
            flags, case.nx, case.ny, case.tau, use_force=True, fx=case.force[0], fy=case.force[1],
            rho0=rho0.astype(np.float64), u0=(u0 if uin is None else _uin(u0, flags, uin)).astype(np.float64),
            alpha=a, threads=min(16, os.cpu_count() or 1))
        t = time.perf_counter()
        Jr, gr, duinr = mk(a64).adjoint(steps, col, t0, t1, ckpt)
        cpu_s = time.perf_counter() - t
        x0, x1, y0, y1 = case.extra["window"]
        win = np.zeros((case.ny, case.nx), bool)
        win[y0:y1 + 1, x0:x1 + 1] = True
        win = win.ravel()
        out["vs_fp64_adjoint"] = {
            "J_rel": abs(J - Jr) / abs(Jr),
            "dJ_dalpha_relL2_window": float(np.linalg.norm(g[win] - gr[win]) / np.linalg.norm(gr[win])),
            "dJ_duin_rel": abs(duin - duinr) / abs(duinr)}
        out["fp64_handle"]["vs_fp64_adjoint"] = {
            "J_rel": abs(J64 - Jr) / abs(Jr),
            "dJ_dalpha_relL2_window": float(np.linalg.norm(g64[win] - gr[win]) / np.linalg.norm(gr[win])),
            "dJ_duin_rel": abs(duin64 - duinr) / abs(duinr)}
        rng = np.random.default_rng(1234)
        idx = rng.choice(np.flatnonzero(win), args.fd_cells, replace=False)
        h = 1e-3
        errs, errs64, errs64h = [], [], []
        for n in idx:
            ap, am = a64.copy(), a64.copy()
            ap[n] += h
            am[n] -= h
            fd = (mk(ap).objective(steps, col, t0, t1) - mk(am).objective(steps, col, t0, t1)) / (2 * h)
            errs.append(abs(g[n] - fd) / abs(fd))
            errs64.append(abs(gr[n] - fd) / abs(fd))
            errs64h.append(abs(g64[n] - fd) / abs(fd))
        uin = float(np.float32(case.u_inlet))
        hu = 3e-4  # inside the range where the difference is stable (1e-3 .. 3e-5)
        fdu = (mk(a64, uin + hu).objective(steps, col, t0, t1)
               - mk(a64, uin - hu).objective(steps, col, t0, t1)) / (2 * hu)
        out["vs_finite_differences"] = {
            "cells": int(args.fd_cells), "h": h, "max_rel_err_gpu": float(max(errs)),
            "median_rel_err_gpu": float(np.median(errs)), "max_rel_err_fp64_adjoint": float(max(errs64)),
            "duin_rel_err_gpu": abs(duin - fdu) / abs(fdu), "duin_rel_err_fp64": abs(duinr - fdu) / abs(fdu),
            "max_rel_err_gpu_fp64_handle": float(max(errs64h)),
            "duin_rel_err_gpu_fp64_handle": abs(duin64 - fdu) / abs(fdu)}
        out["cpu_baseline"] = {"value": round(cpu_s, 3), "unit": "s/gradient",
                               "cores": min(16, os.cpu_count() or 1), "kind": "port",
                               "sample": "FP64 oracle adjoint of the full run with OpenMP threads"}
    print(json.dumps(out), flush=True)


def _uin(u0, flags, v):
    """u0 (as float64) with the inlet cells' x velocity set to v: set in a
    float32 array, u_in +- 1e-5 would round by 4e-9 and the central difference
    be off by 4e-4."""
    u = u0.astype(np.float64)
    u[:flags.size][(flags & 0x20) != 0] = v
    return u


def bench_batch(args):
    """256 single-phase culverts of uneven sizes: one launch per step for all
    (graph replay, autotuned block shape) against launching the instances one
    at a time as the reference's loop does (one launch + device synchronise
    per step, src/cuda/runCuda.cu:473-490)."""
    from flowlb_amd import cases, core
    insts = cases.batch_instances(args.instances)
    built = []
    for c in insts:
        flags, rho0, u0 = c.build()
        built.append(dict(nx=c.nx, ny=c.ny, tau=c.tau, flags=flags, rho0=rho0, u0=u0))
    ncells = sum(c.N for c in insts)
    # Algorithmic bytes of a step: 18 population values + 1 flag byte per fluid
    # cell, 1 flag byte per wall cell.
    active = sum(int(np.count_nonzero((d["flags"] & 0x0A) == 0)) for d in built)
    per_cell = 18 * STORAGE_BYTES[args.storage] + 1
    alg_bytes = active * per_cell + (ncells - active)
    with core.Batch(built, storage=args.storage) as b:
        if args.batch_config:
            # A fixed configuration instead of the autotuner (profiling runs):
            # order,block_x,grid with grid 0 = one block per tile.
            order, bx, grid = args.batch_config.split(",")
            bx, grid = int(bx), int(grid)
            b.set_order(order)
            b.configure(block_x=bx, grid=grid)
        else:
            bx, grid, tune_ms = b.autotune(steps=6)
        b.step(args.warmup)
        b.sync()
        ms = b.step_timed(args.steps)
        info = b.info()
    batched_ms = ms / args.steps
    # one at a time: every instance on its own handle, per-step launch + sync
    per_sync, per_async = 0.0, 0.0
    k = args.single_steps
    for c, d in zip(insts if k > 0 else [], built):
        with core.Solver(d["flags"], c.nx, c.ny, c.options(args.storage), rho0=d["rho0"],
                         u0=d["u0"]) as s:
            s.step(2)
            s.sync()
            t = time.perf_counter()
            for _ in range(k):
                s.step(1)
                s.sync()
            per_sync += (time.perf_counter() - t) / k
            t = time.perf_counter()
            s.step(k)
            s.sync()
            per_async += (time.perf_counter() - t) / k
    out = {"metric": METRIC, "value": round(ncells / (batched_ms * 1e-3) / 1e6, 1), "unit": "MLUPS",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(batched_ms, 5), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None,
           "dtype": "fp32" if args.storage == "fp32" else f"fp32 arithmetic, {args.storage} storage",
           "data": "synthetic",
           "config": {"workload": f"batch_{args.instances}", "instances": args.instances,
                      "storage": args.storage,
                      "cells": ncells, "autotuned_block_x": bx, "fixed_configuration": bool(args.batch_config),
                      "autotuned_mode": "persistent" if grid < 0 else "launch per step",
                      "autotuned_grid": (f"{-grid} resident blocks per CU" if grid < 0 else
                                         ("one block per tile" if grid == 0 else f"{grid} blocks")),
                      "autotuned_tile_order": info["order"],
                      "steps_per_graph_or_cooperative_launch": info["graph_chunk"],
                      "tiles": info["tiles"]},
           "active_fraction": round(active / ncells, 4),
           "achieved_GBs": round(alg_bytes / (batched_ms * 1e-3) / 1e9, 1)}
    traffic = None
    tf = ROOT / "profiles" / f"traffic_batch_{args.instances}_{args.storage}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    out["roofline"] = {"bound": "hbm", "achieved": out["achieved_GBs"], "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(out["achieved_GBs"] / HBM_PEAK_GBS, 4),
                       "traffic": traffic, "kernel": "batched stream-collide",
                       "algorithmic_bytes_per_launch": alg_bytes,
                       "timing": "HIP events around the timed steps (graph replays)"}
    if k > 0:
        out.update({"one_at_a_time_ms_per_step": round(per_sync * 1e3, 4),
                    "one_at_a_time_async_ms_per_step": round(per_async * 1e3, 4),
                    "speedup_vs_one_at_a_time": round(per_sync * 1e3 / batched_ms, 3),
                    "speedup_vs_one_at_a_time_async": round(per_async * 1e3 / batched_ms, 3),
                    "one_at_a_time_async_GBs": round(alg_bytes / (per_async * 1e3 * 1e-3) / 1e9, 1)})
    print(json.dumps(out), flush=True)


def _fs_discharge(g, nx, ny, column):
    """sum of phi rho u_x over the wet (fluid or interface) cells of a column."""
    fl = g["flags"].reshape(ny, nx)[:, column]
    wet = (fl & 0x05) != 0
    phi = g["phi"].reshape(ny, nx)[:, column].astype(np.float64)
    rho = g["rho"].reshape(ny, nx)[:, column].astype(np.float64)
    ux = g["u"][:nx * ny].reshape(ny, nx)[:, column].astype(np.float64)
    return float(np.sum(np.where(wet, phi * rho * ux, 0.0)))


def _front(flags, nx, ny):
    wet = (flags.reshape(ny, nx) & 0x05) != 0
    return int(np.nonzero(wet.any(axis=0))[0].max())


def _surface_mean(phi, nx, ny, k=8):
    """Per column the highest cell with phi >= 0.5 (+ its linear crossing),
    averaged over groups of k columns."""
    p = phi.reshape(ny, nx)
    full = p >= 0.5
    top = np.where(full.any(axis=0), ny - 1 - np.argmax(full[::-1], axis=0), 0)
    h = top.astype(np.float64)
    cols = np.arange(nx)
    above = np.minimum(top + 1, ny - 1)
    a, b = p[top, cols], p[above, cols]
    ok = full.any(axis=0) & (top + 1 < ny) & (a > b)
    h[ok] += (a[ok] - 0.5) / (a[ok] - b[ok])
    return h.reshape(-1, k).mean(axis=1)


def bench_accuracy(args):
    """Storage formats against the FP64 oracle on the full 8192 x 2048 culvert:
    the collapsing column (corrected gas gate) for 3000 steps; time-averaged
    discharge through column 3620 (the front crosses it before step 1000;
    mean over steps 1000-3000, every 100), front position and the surface
    profile (8-column means).  The FP64 oracle runs on the host's cores."""
    from flowlb_amd import cases
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    case = cases.FS_8192
    nx, ny = case.nx, case.ny
    col, t_from, every, steps = 3620, 1000, 100, 3000
    samples = list(range(t_from, steps + 1, every))
    res = {}
    for st in ("fp32", "fp16s", "bf16s"):
        qs = []
        with case.solver(st, mass_mode=1, gas_check_corrected=1) as s:
            t = 0
            for ts in samples:
                s.step(ts - t)
                t = ts
                g = s.read()
                qs.append(_fs_discharge(g, nx, ny, col))
        res[st] = dict(q=float(np.mean(qs)), front=_front(g["flags"], nx, ny),
                       surface=_surface_mean(g["phi"], nx, ny), u=g["u"][:nx * ny])
        print(f"gpu {st}: discharge {res[st]['q']:.6e}, front {res[st]['front']}", flush=True)
    threads = min(16, os.cpu_count() or 1)
    flags, _, _ = case.build()
    o = pyoracle.Oracle(flags, nx, ny, case.tau, use_force=True, fx=case.force[0], fy=case.force[1],
                        free_surface=True, mass_mode=1, gas_check_corrected=True, threads=threads)
    qs, t, t0 = [], 0, time.perf_counter()
    for ts in samples:
        while t < ts:
            o.step(every)
            t += every
            print(f"oracle t={t} ({time.perf_counter() - t0:.0f} s)", flush=True)
        qs.append(o.fs_discharge(col))
    r = o.get()
    o.close()
    q64, front64 = float(np.mean(qs)), _front(r["flags"], nx, ny)
    surf64 = _surface_mean(r["phi"], nx, ny)
    out = {"metric": "storage accuracy vs FP64 oracle", "config": {
        "workload": case.name, "nx": nx, "ny": ny, "steps": steps, "gas_check_corrected": 1,
        "discharge_column": col, "discharge_steps": [t_from, steps, every]},
        "oracle": {"discharge": q64, "front": front64, "threads": threads,
                   "seconds": round(time.perf_counter() - t0, 1)}}
    for st, v in res.items():
        out[st] = {"discharge": v["q"], "discharge_rel_err": abs(v["q"] - q64) / abs(q64),
                   "front": v["front"], "front_err_cells": v["front"] - front64,
                   "surface_mean_abs_err_cells": float(np.abs(v["surface"] - surf64).mean()),
                   "u_x_relL2": float(np.linalg.norm(v["u"] - r["u"][:nx * ny])
                                      / np.linalg.norm(r["u"][:nx * ny]))}
    print(json.dumps(out), flush=True)


def copy_probe(args):
    """A plain device copy of known size (torch, 1 GiB read + 1 GiB written per
    launch): the reference point that scales the FETCH_SIZE / WRITE_SIZE
    counters of the other kernels to bytes."""
    import torch
    n = (1 << 30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                      "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="free_surface_8192x2048")
    ap.add_argument("--storage", default="fp32", choices=list(STORAGE_BYTES))
    ap.add_argument("--profile-steps", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-hbm-reference", action="store_true",
                    help="skip the copy / in-place rates measured in a child process")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--t0", type=int, default=1000,
                    help="lattice time the workload is advanced to before the warm-up, so "
                         "that the timed steps start at t0 + warmup whatever --warmup is")
    ap.add_argument("--no-moving", action="store_true",
                    help="skip the free-surface line under a moving interface")
    ap.add_argument("--fd-cells", type=int, default=16)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--single-steps", type=int, default=20,
                    help="batch: steps per handle of the one-at-a-time comparison (0: skip it)")
    ap.add_argument("--batch-config", default="",
                    help="batch: order,block_x,grid instead of autotuning (e.g. bands,512,0)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the workload is one domain on one MI355X, and "
              "the solver has no multi-GPU path (--gpus 1 only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy_probe":
        return copy_probe(args)
    if args.workload == "hbm_reference":
        print(json.dumps(hbm_reference()), flush=True)
        return
    if args.workload.startswith("adjoint"):
        return bench_adjoint(args)
    if args.workload.startswith("batch"):
        return bench_batch(args)
    if args.workload == "accuracy_8192x2048":
        return bench_accuracy(args)
    # The child runs before this process loads the solver library or touches
    # the GPU -- and not at all under a profiler, whose preloaded library has
    # started the GPU runtime in this process already.
    profiled = "rocprof" in os.environ.get("LD_PRELOAD", "") or any(
        k.startswith("ROCPROF") for k in os.environ)
    hbm_ref = {} if args.no_hbm_reference or profiled else hbm_reference_child()
    from flowlb_amd import cases, core
    case = cases.CASES[args.workload]
    probes = {}
    if not args.no_hbm_reference:
        for kind in core.PROBE_KINDS:
            pms, pbytes = core.bandwidth_probe(case.nx, case.ny, kind, 20)
            probes[f"{kind}_GBs"] = round(pbytes / (pms * 1e-3) / 1e9, 1)
    r = measure(case, args.storage, args.steps, args.warmup, args.profile_steps, t0=args.t0)
    ms_step = r["ms"] / args.steps
    mlups = case.N / (ms_step * 1e-3) / 1e6
    achieved = r["sc_bytes"] / (r["sc_ms"] * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.workload}_{args.storage}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.storage == "fp32" else f"fp32 arithmetic, {args.storage} storage",
        "data": "synthetic",
        "config": {"workload": case.name, "nx": case.nx, "ny": case.ny, "storage": args.storage,
                   "free_surface": case.free_surface, "tau": case.tau,
                   "force": list(case.force), "mass_mode": "conservative",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "stream_collide" + (" (one launch: surface chunks, interior fluid "
                                                   "and bodies of fluid)" if case.free_surface
                                                   else ""),
                     "algorithmic_bytes_per_launch": int(r["sc_bytes"]),
                     "avg_launch_ms": round(r["sc_ms"], 5),
                     "timing": f"HIP events on the solver stream around each stream-collide "
                               f"launch of {args.profile_steps} steps right after the timed "
                               f"region"},
        "lattice_time": [r["t_start"], r["t_start"] + args.steps],
        "active_fraction": round(r["active"] / case.N, 4),
        "mlups_active_cells": round(r["active"] / (ms_step * 1e-3) / 1e6, 1),
        "step_roofline": {"algorithmic_bytes_per_step": int(r["step_bytes"]),
                          "achieved_GBs": round(r["step_bytes"] / (ms_step * 1e-3) / 1e9, 1),
                          "frac": round(r["step_bytes"] / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_ms": {k: round(v, 5) for k, v in zip(core.KERNEL_NAMES, r["per_kernel_ms"])},
        "wall_s_timed_region": round(r["wall"], 4),
    }
    if hbm_ref or probes:
        ref = dict(hbm_ref)
        ref.update(probes)
        if traffic:
            ref["stream_collide_measured_traffic_GBs"] = round(traffic / (r["sc_ms"] * 1e-3) / 1e9, 1)
        # The probe of the same element size: stream-collide's algorithmic rate
        # over what its memory pattern alone attains on this device.
        pat = {"fp32": "esoteric_inplace_GBs", "fp16s": "esoteric_inplace16_GBs",
               "bf16s": "esoteric_inplace16_GBs"}.get(args.storage)
        if pat in probes:
            ref["stream_collide_over_pattern_probe"] = round(achieved / probes[pat], 3)
        out["hbm_reference"] = ref
    if not args.no_sweep:
        sweep = {}
        for st in STORAGE_BYTES:
            if st == args.storage:
                sweep[st] = out["value"]
                continue
            q = measure(case, st, max(50, args.steps // 2), args.warmup // 2, 10, t0=args.t0)
            sweep[st] = round(case.N / (q["ms"] / max(50, args.steps // 2) * 1e-3) / 1e6, 1)
        out["storage_mlups"] = sweep
        out["fp16s_speedup"] = round(sweep["fp16s"] / sweep["fp32"], 3)
    if case.free_surface and not args.no_moving:
        # The reference's gas gate keeps the column almost still; with the
        # corrected gate it collapses: timed after 2000 steps of collapse.
        m = measure(case, args.storage, 200, 0, 10, t0=2000, gas_check_corrected=1)
        mms = m["ms"] / 200
        out["moving_interface"] = {
            "gas_check_corrected": 1, "lattice_time": [m["t_start"], m["t_start"] + 200],
            "MLUPS": round(case.N / (mms * 1e-3) / 1e6, 1),
            "mlups_active_cells": round(m["active"] / (mms * 1e-3) / 1e6, 1),
            "active_fraction": round(m["active"] / case.N, 4), "ms_per_step": round(mms, 5),
            "kernel_ms": {k: round(v, 5) for k, v in zip(core.KERNEL_NAMES, m["per_kernel_ms"])}}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
