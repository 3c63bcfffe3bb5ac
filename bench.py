#!/usr/bin/env python3
"""Benchmark of the FD3D_TSN time-stepping core on one MI355X.

A step is one time step of the plain solver (velocity and stress sweeps,
split-node fault boundary, per-step fault output slabs written to the
device ring) on the single large slip-weakening rupture:
1024x512x512 interior cells + 20-node PML (1064x532x552 = 312.5M cells),
dh 50 m, dt 2 ms, full space, von Karman prestress.  Inputs are synthetic
(seeded fields), resident in HBM before the timed region.

Metric: Gcell-updates/s = nxt*nyt*nzt*steps / time (all cells, PML
included, the SRL 2020 convention), timed on the production path (one
captured step graph replayed per step).  The same number of steps is then
run kernel by kernel with HIP events around every launch: the JSON line
carries the roofline of the dominant kernel (algorithmic bytes per launch
over its event-timed average duration), every sweep against the measured
ceiling of its access mix, and the reference CPU build timed on a bounded
sample on this host.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell of each sweep (FP32 words x 4): the words each
# sweep must move.  The medium is depth-layered (one lam, mu, rho per z
# plane, as the reference builds it), so no sweep streams medium arrays.
BYTES_PER_CELL = {
    "vel_interior": 12 * 4,   # reads 6 stresses + u, v, w; writes u, v, w
    "str_interior": 15 * 4,   # reads u, v, w + 6 stresses; writes 6 stresses
    "pml_vel": 27 * 4,        # 9 splits read+write, 6 stresses read, 3 velocities written
    "pml_str": 39 * 4,        # 15 splits read+write, 3 velocities read, 6 stresses written
    "fused_interior": 18 * 4,  # reads 6 stresses + u, v, w; writes u, v, w + 6 stresses
}
# The same per step: 108 B per bulk cell with the separate sweeps (72 B with
# the fused sweep), 264 B per PML cell.  The survey's convention counts the
# two separate sweeps with the medium as three streamed 3D arrays (120 B /
# 276 B); the JSON line reports the HBM fraction on both.
MOVED_BULK, MOVED_BULK_FUSED, MOVED_PML = 108.0, 72.0, 264.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="large_sw",
                    choices=["large_sw", "plain_sw", "fvw", "batched", "served", "served-cpu", "cpu-reference"])
    ap.add_argument("--batch-mode", default="staged", choices=["persistent", "staged"])
    ap.add_argument("--batch-tz", type=int, default=16)
    ap.add_argument("--batch-fused", default="on", choices=["on", "off"],
                    help="members of a batch ask for the fused interior sweep")
    ap.add_argument("--batch-n", type=int, default=64, help="domains of the batched workload")
    ap.add_argument("--serve-n", type=int, default=48, help="requests of the served stream sample")
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-chunk", type=int, default=16)
    ap.add_argument("--serve-which", default="both", choices=["both", "served", "plain"],
                    help="served workload: the serving loop, the back-to-back plain solves, or both")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="served-cpu workload: concurrent reference processes (at most the host's cores)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 64-domain batched leg of the default line")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="time steps of the reference CPU build on the same grid (about 2 s each on "
                         "1024x512x512: the sample stays within 10-30 s)")
    ap.add_argument("--no-outputs-on", action="store_true",
                    help="skip the pass that streams every step's output slabs to the host")
    return ap.parse_args()


def workload(name):
    from fd3d_tsn_amd import cases as C
    if name == "large_sw":
        return C.c2_large_sw()
    if name == "plain_sw":
        return C.c1_plain_sw()
    return C.c3_fvw()


def cells_of(case, kernel):
    nxi = case.nxt - 2 * case.nabc
    nyi = case.nyt - 2 - case.nabc
    nzi = case.nzt - case.nfs - case.nabc
    if kernel in ("vel_interior", "str_interior", "fused_interior"):
        return nxi * nyi * nzi
    return None


def profiled_traffic(workload_name, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by the
    plain-copy calibration recorded in the same file); None when absent."""
    path = os.path.join(ROOT, "profiles", f"{workload_name}_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) == 3 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel:
                return float(w[2]), os.path.relpath(path, ROOT)
    return None, None


STREAM_MIX = {"vel_interior": "velocity_sweep_mix", "str_interior": "stress_sweep_mix",
              "fused_interior": "fused_sweep_mix", "pml_vel": "pml_velocity_mix", "pml_str": "pml_stress_mix"}


def stream_ceiling(kernel):
    """Measured HBM rate of a stencil-free kernel with this sweep's words
    read and written per cell (profiles/hbm_stream_ceilings.md, made by
    fd3d_tsn_amd/lib/hbm_stream); None when absent."""
    path = os.path.join(ROOT, "profiles", "hbm_stream_ceilings.md")
    if kernel not in STREAM_MIX or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            w = line.split()
            if len(w) == 3 and w[0] == "stream_ceiling_GBs" and w[1] == STREAM_MIX[kernel]:
                return float(w[2])
    return None


def cpu_baseline(case, steps):
    """Reference CPU build (oracle/_ref, flang -O3 -ffast-math, serial) on the
    same grid for a few steps; falls back to the oracle restatement."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refrun
    sample = case.replace(ntfd=steps, stations=[])
    if refrun.available(sample, fast=True):
        res, wall, loop = refrun.run(sample, fast=True, keep_outputs=False, cores="0")
        t = loop if loop else wall
        return {"value": sample.cells * steps / t / 1e9, "unit": "Gcell-updates/s", "cores": 1,
                "kind": "reference",
                "sample": f"{case.name}: {case.nxt}x{case.nyt}x{case.nzt} cells, {steps} time steps, "
                          f"reference Fortran built with flang -O3 -ffast-math, serial, pinned to core 0; "
                          f"time = its own 'CPU TIME OF TIME LOOP' ({t:.2f} s)"}
    import ctypes
    import oracle_lib as O
    O.lib()
    # the oracle's sweeps use OpenMP threads; the baseline is one core
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(1)
    o = O.Oracle(sample, O.F32)
    t0 = time.perf_counter()
    o.step(steps, want_out=False)
    t = time.perf_counter() - t0
    return {"value": sample.cells * steps / t / 1e9, "unit": "Gcell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{case.name}: {steps} steps of the C++ oracle restatement, serial ({t:.2f} s)"}


def bench_cpu_reference(a):
    """The reference CPU build alone on the 1024x512x512 grid for --cpu-steps
    steps, in runs of at most 20 steps (each run prints its time), pinned to
    core 0; the rate comes from the runs' own 'CPU TIME OF TIME LOOP'."""
    from fd3d_tsn_amd import cases as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refrun
    case = C.c2_large_sw()
    model = "unknown"
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    left, loop, runs = a.cpu_steps, 0.0, []
    while left > 0:
        n = min(20, left)
        s = case.replace(ntfd=n, stations=[])
        _, wall, t = refrun.run(s, fast=True, keep_outputs=False, cores="0")
        loop += t
        runs.append({"steps": n, "loop_s": t, "wall_s": wall})
        left -= n
        print(f"[bench] cpu reference: {a.cpu_steps - left} of {a.cpu_steps} steps, this run {t:.2f} s",
              file=sys.stderr, flush=True)
    print(json.dumps({"metric": "Gcell-updates/s (reference CPU build)", "value": case.cells * a.cpu_steps / loop / 1e9,
                      "unit": "Gcell-updates/s", "cores": 1, "kind": "reference", "steps": a.cpu_steps,
                      "ms_per_step": loop / a.cpu_steps * 1e3, "cpu_model": model,
                      "config": {"workload": case.name, "grid": [case.nxt, case.nyt, case.nzt]},
                      "note": "reference Fortran built with flang -O3 -ffast-math, serial, pinned to core 0; "
                              "time = the runs' own 'CPU TIME OF TIME LOOP'", "runs": runs}))


def bench_batched(a, emit=True):
    """Batched uneven domains (64 slip-weakening ruptures, sides 96..320,
    1500 steps each): time per step of the batch against the sum of the
    per-step times of each domain stepped alone (back-to-back plain
    solves); forward ruptures/s follow from the 1500-step length."""
    from fd3d_tsn_amd import autotune
    from fd3d_tsn_amd import capi
    from fd3d_tsn_amd import cases as C
    cases = C.c4_batched(a.batch_n)
    tz, plan = autotune.batch_plan(cases, fused=a.batch_fused == "on")
    tuned = sum(p is not None for p in plan)
    nsteps_run = cases[0].ntfd
    t0 = time.perf_counter()
    solvers = []
    setup_s = []
    phases = {}
    for c in cases:
        t1 = time.perf_counter()
        solvers.append(capi.Solver(c))
        setup_s.append(time.perf_counter() - t1)
        for k, v in solvers[-1].setup_times.items():
            phases[k] = phases.get(k, 0.0) + v / len(cases)
    print(f"[bench] batched: {len(cases)} domains, {sum(c.cells for c in cases)/1e6:.0f}M cells, "
          f"setup {time.perf_counter()-t0:.1f} s", file=sys.stderr, flush=True)
    plain = []
    for s in solvers:
        autotune.apply(s)            # each problem alone gets its own best shape
        s.step(a.warmup, outputs=False)
        s.sync()
        t1 = time.perf_counter()
        s.step(a.steps, outputs=False)
        s.sync()
        plain.append((time.perf_counter() - t1) / a.steps)
    # bytes each member moves per step in the batch (fused members: 72 B per bulk cell)
    # (the batch engine runs the fused stage when the batch holds at least
    # 8 fused work items per CU, as the 64-domain workload does)
    step_b = [((MOVED_BULK_FUSED if p is not None and p[0] & 4 else MOVED_BULK) * nb + MOVED_PML * npml)
              for c, p, (_, nb, npml) in zip(cases, plan, (s.step_bytes() for s in solvers))]
    step_plain = [MOVED_BULK * nb + MOVED_PML * npml for _, nb, npml in (s.step_bytes() for s in solvers)]
    for s, p in zip(solvers, plan):
        s.init()
        if p is not None:
            s.set_tuning(*p)
    mode = capi.PERSISTENT if a.batch_mode == "persistent" else capi.STAGED
    if not tuned:
        tz = a.batch_tz
    b = capi.Batch(len(solvers), tz=tz, mode=mode)
    for s in solvers:
        b.add(s)
    b.step(a.warmup)
    b.sync()
    b.elapsed()
    t1 = time.perf_counter()
    b.step(a.steps)
    b.sync()
    wall = (time.perf_counter() - t1) / a.steps
    gpu_ms, n = b.elapsed()
    info = b.info()
    b.close()
    for s in solvers:
        s.close()
    cells = sum(c.cells for c in cases)
    t_plain = sum(plain)
    step_bytes = sum(step_b)
    value = len(cases) / (nsteps_run * wall)
    out = {"metric": "forward ruptures/s (batched uneven domains)", "value": value,
           "unit": "ruptures/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"batched_{len(cases)}_domains_96-320", "domains": len(cases),
                      "cells": cells, "steps_per_rupture": nsteps_run, "mode": a.batch_mode,
                      "tz": tz, "autotuned_members": tuned,
                      "fused_members": sum(1 for c, p in zip(cases, plan) if p is not None and p[0] & 4)},
           "batch": {"gpu_ms_per_step": gpu_ms / max(n, 1), "Gcell_updates_per_s": cells / wall / 1e9,
                     "hbm_frac": step_bytes / wall / 1e9 / HBM_PEAK_GBS, "info": info},
           "plain_back_to_back": {"ms_per_step_sum": t_plain * 1e3,
                                  "ruptures_per_s": len(cases) / (nsteps_run * t_plain),
                                  "Gcell_updates_per_s": cells / t_plain / 1e9,
                                  "hbm_frac": sum(step_plain) / t_plain / 1e9 / HBM_PEAK_GBS,
                                  "setup_s_per_rupture": sum(setup_s) / len(setup_s),
                                  "setup_phases_s": phases},
           "speedup_vs_plain": t_plain / wall}
    if emit:
        print(json.dumps(out))
    return out


def bench_served(a):
    """Serving loop on the first --serve-n requests of the Poisson stream
    (mean 20 requests/s, 70% slip weakening / 30% rate-and-state, grids
    64^3..384x192x192, 800-3000 steps): continuous batching against the same
    requests solved back to back; per-request relative L2 of the final
    arrays between the two."""
    from fd3d_tsn_amd import capi
    from fd3d_tsn_amd import serve
    reqs = serve.c5_requests(2000)[:a.serve_n]
    upd = sum(r.case.cells * r.steps for r in reqs)
    print(f"[bench] served: {len(reqs)} requests, {upd/1e9:.0f} G cell-updates", file=sys.stderr, flush=True)
    # warm the library and the device once
    with capi.Solver(reqs[0].case.replace(nxtT=32, nytT=16, nztT=16, ntfd=4)) as s:
        s.step(4, outputs=False)
        s.sync()
    def progress(tag, every=25):
        n = [0]

        def log(msg):
            if "done" in msg or "failed" in msg or tag == "plain":
                n[0] += 1
                if n[0] % every == 0 or n[0] == len(reqs):
                    print(f"[bench] {tag}: {n[0]} of {len(reqs)} requests finished", file=sys.stderr, flush=True)
        return log

    plain, got, srv = None, None, None
    keep = a.serve_n <= 200     # final arrays of both legs for the per-request comparison
    if a.serve_which in ("both", "plain"):
        t1 = time.perf_counter()
        plain = serve.run_plain(reqs, keep_final=keep, log=progress("plain"))
        print(f"[bench] plain back-to-back {time.perf_counter()-t1:.1f} s", file=sys.stderr, flush=True)
    if a.serve_which in ("both", "served"):
        srv = serve.Server(max_slots=a.serve_slots, chunk=a.serve_chunk, tz=a.batch_tz, fused=a.batch_fused == "on",
                           keep_final=keep, log=progress("served"))
        t1 = time.perf_counter()
        got = srv.run(reqs)
        print(f"[bench] served {time.perf_counter()-t1:.1f} s", file=sys.stderr, flush=True)
        srv.close()
    ps = serve.summary(plain, max(r.finished for r in plain)) if plain else None
    ss = serve.summary(got, srv.makespan) if got else None
    l2 = None
    if plain and got and keep:
        pmap = {r.rid: r for r in plain}
        l2 = 0.0
        for r in got:
            for k, v in pmap[r.rid].final.items():
                l2 = max(l2, serve.rel_l2(r.final[k], v))
    lead = ss or ps
    out = {"metric": "served ruptures/s (Poisson stream, continuous batching)" if ss else
                     "back-to-back plain ruptures/s (Poisson stream)", "value": lead["ruptures_per_s"],
           "unit": "ruptures/s", "n_gpus": 1, "steps": None, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (rate-and-state Newton and state in f64)", "data": "synthetic",
           "config": {"workload": f"served_stream_first_{len(reqs)}_of_2000", "requests": len(reqs),
                      "rate_per_s": 20.0, "slots": a.serve_slots, "chunk_steps": a.serve_chunk,
                      "cell_updates": upd},
           "served": dict(ss, gpu_step_s=srv.gpu_step_s, host_wait_s=srv.gpu_busy_s, chunks=srv.chunks,
                          host_s=srv.host_s) if ss else None,
           "plain_back_to_back": ps,
           "speedup_vs_plain": ss["ruptures_per_s"] / ps["ruptures_per_s"] if ss and ps else None,
           "max_rel_l2_vs_plain": l2}
    print(json.dumps(out))


def bench_served_cpu(a):
    """The reference CPU build on the served mix: the first --serve-n
    requests of the Poisson stream (full grids and step counts), run by
    --cpu-procs concurrent reference processes, each pinned to its own host
    core (the reference is serial); aggregate ruptures/s and cell-updates/s.
    The mix's rate follows from the cell-update throughput and the mean work
    per request of the whole 2000-request stream."""
    import queue
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import refrun
    from fd3d_tsn_amd import serve
    allreq = serve.c5_requests(2000)
    reqs = allreq[:a.serve_n]
    ncpu = os.cpu_count() or 1
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    nproc = max(1, min(a.cpu_procs, ncpu, len(reqs)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    cores = sorted(os.sched_getaffinity(0))[:nproc] if hasattr(os, "sched_getaffinity") else list(range(nproc))
    todo = queue.Queue()
    for r in reqs:
        todo.put(r)
    done, errors = [], []
    lock = threading.Lock()
    t0 = time.perf_counter()

    def worker(core):
        while True:
            try:
                r = todo.get_nowait()
            except queue.Empty:
                return
            t1 = time.perf_counter()
            try:
                _res, wall, loop = refrun.run(r.case, fast=True, keep_outputs=False, cores=str(core),
                                              timeout=3600)
            except Exception as e:  # reported in the JSON line
                with lock:
                    errors.append(f"request {r.rid}: {str(e)[:200]}")
                continue
            with lock:
                done.append((r.rid, time.perf_counter() - t1, loop))
                print(f"[bench] cpu mix: request {r.rid} ({r.case.friction}, {r.case.cells/1e6:.1f}M cells, "
                      f"{r.steps} steps) {wall:.1f} s on core {core}; {len(done)} of {len(reqs)} done",
                      file=sys.stderr, flush=True)

    threads = [threading.Thread(target=worker, args=(c,)) for c in cores]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    upd = sum(r.case.cells * r.steps for r in reqs if any(d[0] == r.rid for d in done))
    mean_all = sum(r.case.cells * r.steps for r in allreq) / len(allreq)
    gups = upd / wall / 1e9
    out = {"metric": "served ruptures/s (reference CPU build, concurrent processes)",
           "value": len(done) / wall, "unit": "ruptures/s", "n_gpus": 0, "steps": None, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (rate-and-state: the f64 build)", "data": "synthetic",
           "config": {"workload": f"served_stream_first_{len(reqs)}_of_2000_cpu", "requests": len(reqs),
                      "processes": nproc, "cores": cores, "core_model": model, "host_cores": ncpu,
                      "build": "reference Fortran, flang -O3 -ffast-math (-fdefault-real-8 for rate-and-state), "
                               "serial, one process per core (taskset)"},
           "wall_s": wall, "Gcell_updates_per_s": gups,
           "mix_ruptures_per_s": gups * 1e9 / mean_all,
           "mix_note": "the whole stream's rate at this cell-update throughput (mean work per request of "
                       "the 2000 requests)",
           "per_request_s": sorted(d[1] for d in done), "errors": errors}
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "batched":
        return bench_batched(a)
    if a.workload == "served":
        return bench_served(a)
    if a.workload == "served-cpu":
        return bench_served_cpu(a)
    if a.workload == "cpu-reference":
        return bench_cpu_reference(a)
    import numpy as np
    from fd3d_tsn_amd import capi
    from fd3d_tsn_amd import cases as C

    case = workload(a.workload)
    # CPU baseline first, before this process initialises the GPU
    cpu = None
    if not a.no_cpu_baseline:
        print(f"[bench] cpu baseline: {a.cpu_steps} steps of the reference build", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(case, a.cpu_steps)
        except Exception as e:  # reported, not fatal: the GPU number stands
            cpu = {"error": str(e)[:300]}
    print(f"[bench] {case.name}: {case.nxt}x{case.nyt}x{case.nzt} = {case.cells/1e6:.1f}M cells, "
          f"warmup {a.warmup}, steps {a.steps}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    fault, scal = (C.fault_fvw(case) if case.friction == "fvw" else (None, None))
    s = capi.Solver(case, fault=fault, fvw_scalars=scal)
    from fd3d_tsn_amd import autotune
    tuned = autotune.apply(s)
    print(f"[bench] setup {time.perf_counter()-t0:.1f} s", file=sys.stderr, flush=True)
    s.step(a.warmup, outputs=False)
    s.sync()
    # timed region: the production path (one captured step graph replayed
    # per step)
    t0 = time.perf_counter()
    s.step(a.steps, outputs=False)
    s.sync()
    elapsed = time.perf_counter() - t0
    # the same number of steps kernel by kernel, HIP events around every
    # launch on the solver's stream: the per-kernel averages of the roofline
    s.set_profiling(True)
    t0 = time.perf_counter()
    s.step(a.steps, outputs=False)
    s.sync()
    elapsed_prof = time.perf_counter() - t0
    kt = s.kernel_times()
    s.set_profiling(False)
    # the production driver's path: the same number of steps with every
    # step's output slabs and station samples streamed to pinned host
    # memory (fd3d_step_begin / fd3d_step_wait), the host reading each chunk
    # while the GPU computes the next
    outputs_on = None
    if not a.no_outputs_on:
        # at least 64 steps in chunks of 16: the last chunk's copy (the only
        # one not hidden behind computing) is then a quarter of a chunk run
        n_on, chunk = max(a.steps, 64), 16
        s.step(2, outputs=False)
        s.sync()
        t0 = time.perf_counter()
        peak = 0.0
        for res in s.stream(n_on, chunk, copy=False):
            peak = max(peak, float(np.max(np.abs(res["sliprateX"][-1]))))
        el = time.perf_counter() - t0
        nxT, nzT = case.fault_shape
        outputs_on = {"ms_per_step": el * 1e3 / n_on, "steps": n_on, "chunk_steps": chunk,
                      "host_bytes_per_step": 5 * nxT * nzT * 4 + 3 * 4 * len(case.stations),
                      "of_outputs_off": (el / n_on) / (elapsed / a.steps),
                      "note": "every step's 5 output slabs and the station samples copied to pinned host "
                              "memory on a copy stream (fd3d_step_begin / fd3d_step_wait), the host reading "
                              "each chunk while the next computes; timed until the last chunk is on the host"}
    step_bytes, nbulk, npml = s.step_bytes()
    tuning = s.get_tuning()
    fused = bool(tuning[0] & 4)
    moved = (MOVED_BULK_FUSED if fused else MOVED_BULK) * nbulk + MOVED_PML * npml
    s.close()

    ms_per_step = elapsed * 1e3 / a.steps
    value = case.cells * a.steps / elapsed / 1e9
    # dominant kernel by total time
    dom = max(kt.items(), key=lambda kv: kv[1][0])
    name, (ms_tot, launches) = dom
    avg_ms = ms_tot / launches
    ncell = npml if name.startswith("pml") else cells_of(case, name)
    if name in BYTES_PER_CELL and ncell:
        alg = BYTES_PER_CELL[name] * ncell
    else:
        alg = None
    traffic, traffic_src = profiled_traffic(case.name, name)
    roof = {"bound": "hbm", "kernel": name,
            "achieved": (alg / (avg_ms * 1e-3) / 1e9) if alg else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (alg / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if alg else None,
            "traffic": traffic,
            "avg_launch_ms": avg_ms, "alg_bytes_per_launch": alg}
    if traffic_src:
        roof["traffic_source"] = traffic_src
    ceil = stream_ceiling(name)
    if ceil and alg:
        roof["stream_ceiling"] = {"GBs": ceil, "frac": alg / (avg_ms * 1e-3) / 1e9 / ceil,
                                  "source": "profiles/hbm_stream_ceilings.md"}
    kernels = {}
    for k, v in sorted(kt.items()):
        kernels[k] = {"avg_ms": v[0] / v[1], "launches": v[1]}
        ncell = npml if k.startswith("pml") else cells_of(case, k)
        if k in BYTES_PER_CELL and ncell:
            gbs = BYTES_PER_CELL[k] * ncell / (v[0] / v[1] * 1e-3) / 1e9
            kernels[k]["alg_GBs"] = gbs
            c = stream_ceiling(k)
            if c:
                kernels[k]["of_stream_ceiling"] = gbs / c
    out = {
        "metric": "Gcell-updates/s", "value": value, "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": case.name, "grid": [case.nxt, case.nyt, case.nzt], "dh": case.dh,
                   "dt": case.dt, "nabc": case.nabc, "friction": case.friction,
                   "fspace": case.fspace, "cells": case.cells, "tuning": list(tuning),
                   "autotuned": tuned},
        "profiled_ms_per_step": elapsed_prof * 1e3 / a.steps,
        "hbm_step": {"alg_bytes_per_step": moved, "bulk_cells": nbulk, "pml_cells": npml,
                     "achieved_GBs": moved / (ms_per_step * 1e-3) / 1e9,
                     "frac": moved / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "convention_bytes_per_step": step_bytes,
                     "convention_frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "note": f"alg bytes: {72 if fused else 108} B/bulk cell ({'fused' if fused else 'separate'} "
                             "interior sweep), 264 B/PML cell (layered medium, no medium arrays streamed); "
                             "convention: 120/276 B, separate sweeps with the medium as three 3D arrays"},
        "roofline": roof,
        "kernels": kernels,
    }
    if outputs_on is not None:
        out["ms_per_step_outputs_on"] = outputs_on["ms_per_step"]
        out["outputs_on"] = outputs_on
    # the other single-rupture workloads beside it (same steps and warm-up):
    # the rate-and-state rupture and the small parity rupture
    if not a.no_batched and a.workload == "large_sw":
        for wname, wcase in (("fvw_512x256x256", C.c3_fvw()), ("plain_sw_128x64x64", C.c1_plain_sw())):
            try:
                fault2, scal2 = (C.fault_fvw(wcase) if wcase.friction == "fvw" else (None, None))
                with capi.Solver(wcase, fault=fault2, fvw_scalars=scal2) as s2:
                    autotune.apply(s2)
                    s2.step(a.warmup, outputs=False)
                    s2.sync()
                    t0 = time.perf_counter()
                    s2.step(a.steps, outputs=False)
                    s2.sync()
                    el2 = (time.perf_counter() - t0) / a.steps
                    b2, nb2, np2 = s2.step_bytes()
                    t2 = s2.get_tuning()
                moved2 = (MOVED_BULK_FUSED if t2[0] & 4 else MOVED_BULK) * nb2 + MOVED_PML * np2
                out[wname] = {"ms_per_step": el2 * 1e3, "Gcell_updates_per_s": wcase.cells / el2 / 1e9,
                              "hbm_frac_moved_bytes": moved2 / el2 / 1e9 / HBM_PEAK_GBS,
                              "hbm_frac_convention": b2 / el2 / 1e9 / HBM_PEAK_GBS,
                              "grid": [wcase.nxt, wcase.nyt, wcase.nzt], "friction": wcase.friction,
                              "fspace": wcase.fspace, "tuning": list(t2)}
            except Exception as e:  # reported, not fatal
                out[wname] = {"error": str(e)[:300]}
    # the batched workload (64 uneven domains) beside it, same steps and warm-up
    if not a.no_batched and a.workload == "large_sw":
        print("[bench] batched leg: 64 uneven domains", file=sys.stderr, flush=True)
        try:
            bt = bench_batched(a, emit=False)
            out["batched_64_domains"] = {
                "ruptures_per_s": bt["value"], "ms_per_step": bt["ms_per_step"],
                "Gcell_updates_per_s": bt["batch"]["Gcell_updates_per_s"], "hbm_frac": bt["batch"]["hbm_frac"],
                "plain_back_to_back_ruptures_per_s": bt["plain_back_to_back"]["ruptures_per_s"],
                "plain_back_to_back_ms_per_step_sum": bt["plain_back_to_back"]["ms_per_step_sum"],
                "speedup_vs_plain": bt["speedup_vs_plain"], "cells": bt["config"]["cells"],
                "steps_per_rupture": bt["config"]["steps_per_rupture"], "tz": bt["config"]["tz"],
                "fused_members": bt["config"]["fused_members"]}
        except Exception as e:  # reported, not fatal: the main number stands
            out["batched_64_domains"] = {"error": str(e)[:300]}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out))


if __name__ == "__main__":
    main()
