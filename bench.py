#!/usr/bin/env python3
"""Throughput of Pearl's per-step solver loop on one MI355X (FP64 plain form).

Workload (production FP64 domain): 4096 x 2048 cells, the default 10 K
phase-diagram input's physics, a colony of 32 ferrite/cementite lamella
pairs (spacing +/-5 % from a fixed seed, 1:7 cementite:ferrite, colony 256
rows high) over eutectoid-composition austenite, adaptive dt, implicit
diffusion, moving window checked at output times, 10,000 accepted steps
timed after 5 warm-up steps (the defaults).  A "step" is one accepted
step of the host program's loop (all its attempts, PCG iterations and
acceptance reductions): pearl_run between output times, with the reference
loop's output tolerance, then the moving-window check at each output time.
--host-loop advance times one pearl_advance call per step instead.

Prints one JSON line:
  metric  Mcell-steps/s = nx*ny*accepted steps / wall seconds of the timed loop / 1e6
  roofline  the dominant kernel's algorithmic HBM bytes per launch / its
            average launch time, against 8 TB/s.  The launch times come from
            HIP events on the library's stream around each cell kernel, over
            a second region of --kernel-steps steps (500) right after the
            timed one (the events add graph nodes and slow the whole loop by
            ~15 %, so the timed region runs without them); `traffic` is the
            kernel's HBM bytes per launch: its algorithmic bytes per launch
            here times the measured traffic per algorithmic byte (rocprofv3
            FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy kernel) from the
            committed counter summary of the same workload and tile height,
            when there is one, and traffic_GBps that traffic over the live
            launch time;
            copy_ceiling_GBps is a hand-written 1 GiB device copy measured in
            the same run (pearl_stream_probe), frac_of_copy the dominant
            kernel's rate against it; for every cell kernel,
            bare_pattern_GBps is a bare kernel with the same read and write
            planes in the same geometry (pearl_stream_mix) and
            bare_pattern_over_kernel_time its launch time over the kernel's
            (near 1: the kernel runs at its access pattern's rate)
  cpu_baseline  the reference solver compiled from source (or the oracle
            restatement when that is absent) on a bounded sample of the same
            workload on this host's cores
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "fe-c_eutectoid_phase-field_simulation_amd")
HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def load_pkg():
    spec = importlib.util.spec_from_file_location("pearl_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pearl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(pkg, nx, ny, pairs, strip_rows, seed, undercooling, tmpdir, overrides=None):
    import importlib
    inputs = importlib.import_module("pearl_amd.inputs")
    return inputs.colony_workload(nx, ny, pairs, strip_rows, seed, undercooling, tmpdir, overrides=overrides)


def cpu_baseline(path, cfg, phi, c, mu, initial_carbon, seconds, tmpdir):
    """Time the reference on the same state, in child processes (the
    reference library carries its own HDF5 and C++ runtime and is never
    loaded next to the GPU runtime):

    - serial: the reference advance() (built from pearl_serial.cpp) for
      about `seconds`;
    - MPI: the unmodified pearl_parallel under mpiexec on min(16, host CPUs)
      ranks, restarted from the same state written as a reference
      checkpoint, for about the same wall time (step count from the serial
      rate).

    Returns (primary, serial): the MPI figure when it ran, else the serial
    one.  Falls back to the oracle restatement when the reference build is
    absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    import ref_tool
    n = cfg.nx * cfg.ny
    sc = [0.0, cfg.dt, 0.0, initial_carbon, 0.0, 0.0, 0, 0, 0, cfg.initial_mu]
    if O.ref_available():
        state = os.path.join(tmpdir, "baseline_state.npz")
        np.savez(state, phi=phi, carbon=c, mu=mu, scalars=np.array(sc), stop_time=cfg.end_time)
        out = os.path.join(tmpdir, "baseline.json")
        ref_tool.call("baseline", path, state, seconds, out, timeout=int(seconds * 20 + 120))
        r = json.load(open(out))
        steps, el, kind = r["steps"], r["seconds"], "reference"
    else:
        log("cpu baseline: reference build absent; timing the oracle restatement")
        P = O.params_from_dict(cfg.solver_params())
        S = O.OracleState(phi, c, mu, sc[:9])
        steps, kind = 0, "port"
        t0 = time.perf_counter()
        while True:
            rc, _ = O.advance(P, S, cfg.initial_mu, cfg.end_time)
            if rc:
                raise RuntimeError(f"oracle advance failed {rc}")
            steps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                break
    serial = dict(value=n * steps / el / 1e6, unit="Mcell-steps/s", cores=1, kind=kind,
                  sample=f"{steps} accepted steps of the same {cfg.nx}x{cfg.ny} workload from the same initial "
                         f"state, serial ({'pearl_serial.cpp advance' if kind == 'reference' else 'oracle'}), "
                         f"{el:.1f} s")
    if kind != "reference" or not ref_tool.mpiexec() or not ref_tool.parallel_program():
        return serial, serial, None
    nprocs = max(1, min(16, os.cpu_count() or 1))
    while nprocs > 1:
        try:
            ref_tool.grid_split(nprocs, cfg.nx, cfg.ny)
            break
        except ValueError:
            nprocs -= 1
    mpi_steps = max(4, int(steps / el * seconds * nprocs * 0.5))
    out = os.path.join(tmpdir, "baseline_mpi.json")
    work = os.path.join(tmpdir, "mpi")
    try:
        ref_tool.call("baseline-mpi", path, state, nprocs, mpi_steps, out, work, timeout=int(seconds * 40 + 300))
    except Exception as e:  # noqa: BLE001
        log(f"cpu baseline: pearl_parallel run failed ({e}); reporting the serial figure")
        return serial, serial, None
    r = json.load(open(out))
    mpi = dict(value=n * r["steps"] / r["seconds"] / 1e6, unit="Mcell-steps/s", cores=nprocs, kind="reference",
               sample=f"{r['steps']} accepted steps of the same {cfg.nx}x{cfg.ny} workload, restarted from the same "
                      f"state, unmodified pearl_parallel under mpiexec -n {nprocs} ({r['px']}x{r['py']} ranks), "
                      f"step-loop wall {r['seconds']:.1f} s from its diagnostics.csv")
    return mpi, serial, parity_with_reference_run(path, work, mpi_steps)


def parity_with_reference_run(path, work, steps):
    """The drop-in program pearl_gpu restarted from the same reference
    checkpoint for the same number of steps as the pearl_parallel run; both
    final checkpoints read back through the reference reader."""
    import subprocess
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_tool
    exe = os.path.join(PKG_DIR, "pearl_gpu")
    gdir = os.path.join(work, "gpu")
    r = subprocess.run([exe, path, gdir, "--restart", os.path.join(work, "start.h5"), "--max-steps", str(steps)],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"parity run failed: {r.stderr.strip()[-300:]}")
        return None

    def status(d):
        return dict(l.split("=", 1) for l in open(os.path.join(d, "STOPPED.txt")).read().splitlines() if "=" in l)

    import hashlib
    sr, sg = status(os.path.join(work, "run")), status(gdir)
    a = ref_tool.read_checkpoint(path, os.path.join(work, "run", "checkpoint.h5"), os.path.join(work, "run"))
    b = ref_tool.read_checkpoint(path, os.path.join(gdir, "checkpoint.h5"), gdir)
    s0 = ref_tool.read_checkpoint(path, os.path.join(work, "start.h5"), work)
    out = {"what": f"pearl_gpu and pearl_parallel (MPI) restarted from the same reference checkpoint, {steps} steps each",
           "steps": [int(sr["step"]), int(sg["step"])], "rejected": [int(sr["rejected"]), int(sg["rejected"])],
           "cg_total": [int(sr["CG_total"]), int(sg["CG_total"])],
           "max_abs_dmu_J_per_mol": float(np.max(np.abs(a[2] - b[2]))),
           "max_abs_dphi": float(np.max(np.abs(a[0] - b[0]))),
           "max_abs_dc": float(np.max(np.abs(a[1] - b[1]))),
           # how far the runs moved from the common start (so equal fields are not a stale read)
           "moved_from_start_max_abs_dphi": float(np.max(np.abs(a[0] - s0[0]))),
           "moved_from_start_max_abs_dmu": float(np.max(np.abs(a[2] - s0[2]))),
           "mu_sha1": [hashlib.sha1(a[2].tobytes()).hexdigest()[:12], hashlib.sha1(b[2].tobytes()).hexdigest()[:12]]}
    log(f"parity at {a[2].shape[1]}x{a[2].shape[0]}: steps {out['steps']}, rejected {out['rejected']}, CG "
        f"{out['cg_total']}, max|dmu| {out['max_abs_dmu_J_per_mol']:.2e}, max|dphi| {out['max_abs_dphi']:.2e}")
    return out


def build_ensemble(pkg, count, seed, tmpdir):
    import importlib
    return importlib.import_module("pearl_amd.inputs").ensemble_workload(count, seed, tmpdir)


def run_mixed_accuracy(args, pkg):
    """FP32-mixed form against the FP64 plain form on the same inputs.
    --acc-mode lockstep: each step the FP64 context advances, then the mixed
    context takes a step of exactly the accepted dt (so both follow one dt
    sequence and the difference is precision alone); the fields are
    compared every --acc-every steps.  --acc-mode free: each form runs its
    own adaptive dt and the fields are compared at matched physical times,
    about every --acc-every FP64 steps.  Compared: max |dmu| in J/mol and
    max |dphi| per phase fraction.  Afterwards both forms run free from
    their states for --steps steps, the timed throughput comparison."""
    import numpy as np
    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        torch, have_torch = None, False
    nx, ny = (8192, 4096) if (args.nx, args.ny) == (4096, 2048) else (args.nx, args.ny)
    pairs = 64 if args.pairs == 32 else args.pairs
    uc = 20.0 if args.undercooling == 10.0 else args.undercooling
    tmp = tempfile.mkdtemp()
    path, cfg, phi, c, mu = build_workload(pkg, nx, ny, pairs, ny // 8, args.seed, uc, tmp)
    log(f"mixed accuracy ({args.acc_mode}): {nx}x{ny}, {pairs} pairs, {uc} K, {args.acc_steps} FP64 steps, "
        f"compared every {args.acc_every}")
    a = pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=args.cells_per_thread, precision=pkg.FP64)
    if args.acc_other in ("mixed", "stencil32"):
        b = pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=args.cells_per_thread,
                       precision=pkg.MIXED if args.acc_other == "mixed" else pkg.FP32_STENCIL)
    else:
        b = pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=8, precision=pkg.FP64)
    for s in (a, b):
        s.set_state(phi, c, mu, [0.0, cfg.dt, 0.0, 0.0, 0.0, 0.0, 0, 0, 0])
    m0 = a.total_carbon()
    rows, broken = [], None
    if args.acc_mode == "free":
        # Each form runs its own adaptive dt; they are compared at matched
        # physical times t_k = k T, where T is the time the FP64 form reaches
        # in its first --acc-every steps.  pearl_run clips the last step of
        # each interval to t_k, so both states are taken exactly there.
        big = 10 ** 8
        T = None
        for k in range(1, max(1, args.acc_steps // args.acc_every) + 1):
            sa0, sb0 = a.scalars().step, b.scalars().step
            if T is None:
                a.run(1.0, args.acc_every)
                T = a.scalars().time
                t_k = T
            else:
                t_k = k * T
                a.run(t_k, big)
            b.run(t_k, big)
            ta, tb = a.scalars().time, b.scalars().time
            pa, ca, ma, _ = a.get_state()
            pb, cb, mb, _ = b.get_state()
            row = {"interval": k, "time": ta, "time_other": tb,
                   "steps": [a.scalars().step, b.scalars().step],
                   "interval_steps": [a.scalars().step - sa0, b.scalars().step - sb0],
                   "max_dmu_J_per_mol": float(np.max(np.abs(ma - mb))),
                   "max_dphi": [float(np.max(np.abs(pa[i] - pb[i]))) for i in range(3)],
                   "max_dc": float(np.max(np.abs(ca - cb))),
                   "mass_rel_mixed": abs(b.total_carbon() - m0) / abs(m0)}
            rows.append(row)
            log(f"  t={ta:.4e} (other {tb:.4e}) steps {row['steps']}: max|dmu| {row['max_dmu_J_per_mol']:.3e} "
                f"J/mol, max|dphi| {row['max_dphi']}")
    for k in range(1, args.acc_steps + 1 if args.acc_mode == "lockstep" else 1):
        info = a.advance(1.0)
        sc = b.scalars()
        sc.dt = info.accepted_dt
        b.set_scalars(sc)
        r0 = b.counters()["rejected_attempts"]
        b.advance(1.0)
        if broken is None and (b.counters()["rejected_attempts"] != r0 or b.scalars().time != a.scalars().time):
            broken = k
        if k % args.acc_every == 0 or k == args.acc_steps:
            pa, ca, ma, _ = a.get_state()
            pb, cb, mb, _ = b.get_state()
            row = {"step": k, "time": a.scalars().time, "max_dmu_J_per_mol": float(np.max(np.abs(ma - mb))),
                   "max_dphi": [float(np.max(np.abs(pa[i] - pb[i]))) for i in range(3)],
                   "max_dc": float(np.max(np.abs(ca - cb))),
                   "mass_rel_mixed": abs(b.total_carbon() - m0) / abs(m0)}
            rows.append(row)
            log(f"  step {k}: t={row['time']:.4e} max|dmu| {row['max_dmu_J_per_mol']:.3e} J/mol, "
                f"max|dphi| {row['max_dphi']}")
    # free-running throughput of each form from its own state
    timing = {}
    for name, s in (("fp64", a), ("mixed" if args.acc_other == "mixed" else "other", b)):
        s.run(1.0, 5)
        c0 = s.counters()
        if have_torch:
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = s.run(1.0, args.steps)
        wall = time.perf_counter() - t0
        c1 = s.counters()
        timing[name] = {"value": nx * ny * r["taken"] / wall / 1e6, "steps": r["taken"], "wall_s": wall,
                        "attempts": c1["attempts"] - c0["attempts"],
                        "cg_iterations": c1["cg_iterations"] - c0["cg_iterations"]}
        log(f"  free run {name}: {r['taken']} steps in {wall:.3f} s -> {timing[name]['value']:.1f} Mcell-steps/s")
    a.close()
    b.close()
    other = "mixed" if args.acc_other == "mixed" else "other"
    vmx, v64 = timing[other]["value"], timing["fp64"]["value"]
    out = {
        "metric": "Mcell-steps/s", "value": round(vmx, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": timing[other]["steps"], "warmup": 5,
        "ms_per_step": round(timing[other]["wall_s"] * 1e3 / max(timing[other]["steps"], 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{'FP32-mixed form' if args.acc_other == 'mixed' else 'FP64 with 8-row tiles'} "
                               f"against FP64 on {nx}x{ny}, {pairs} pairs +/-5% (seed {args.seed}), "
                               f"{uc} K; errors "
                               + ("in lockstep (one dt sequence)" if args.acc_mode == "lockstep"
                                  else "at matched physical times, each form on its own dt sequence"),
                   "nx": nx, "ny": ny, "precision": args.acc_other},
        "fp64_value": round(v64, 3), "mixed_over_fp64": round(vmx / v64, 3), "timing": timing,
        "acc_mode": args.acc_mode, "lockstep_steps": args.acc_steps if args.acc_mode == "lockstep" else None,
        "lockstep_broken_at": broken, "comparisons": rows,
        "max_dmu_J_per_mol": max(r["max_dmu_J_per_mol"] for r in rows),
        "max_dphi": [max(r["max_dphi"][i] for r in rows) for i in range(3)],
    }
    print(json.dumps(out), flush=True)


def run_ensemble(args, pkg):
    """Batched ensemble in one context versus the same domains one at a time."""
    import numpy as np
    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        torch, have_torch = None, False
    tmp = tempfile.mkdtemp()
    doms = build_ensemble(pkg, args.domains, args.seed, tmp)
    cells = sum(d["cfg"].nx * d["cfg"].ny for d in doms)
    log(f"ensemble: {len(doms)} domains, {cells / 1e6:.2f} M cells")
    precision = pkg.MIXED if args.precision == "mixed" else pkg.FP64
    params = [d["cfg"].solver_params() for d in doms]
    mus = [d["cfg"].initial_mu for d in doms]
    states = [(d["phi"], d["c"], d["mu"], [0.0, d["cfg"].dt, 0, 0, 0, 0, 0, 0, 0]) for d in doms]
    k = args.cells_per_thread
    tuned = None
    if k == 0:
        import importlib
        autotune = importlib.import_module("pearl_amd.autotune")
        log("autotuning the tile height ...")
        k, tuned = autotune.tune_batch(pkg, params, mus, states, precision=precision, log=log)
        log(f"autotune: {k} rows per tile")
    b = pkg.Batch(params, mus, cells_per_thread=k, precision=precision)
    for i, d in enumerate(doms):
        b.set_domain_state(i, d["phi"], d["c"], d["mu"], [0.0, d["cfg"].dt, 0, 0, 0, 0, 0, 0, 0])
    stops = [1e-6] * len(doms)
    r = b.run_all(stops, [args.warmup] * len(doms))
    assert r["rc"] == 0, r
    if have_torch:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = b.run_all(stops, [args.steps] * len(doms))
    wall = time.perf_counter() - t0
    assert r["rc"] == 0, r
    batched = cells * args.steps / wall / 1e6
    ctr = b.counters()
    b.close()
    log(f"batched: {args.steps} steps per domain in {wall:.3f} s -> {batched:.1f} Mcell-steps/s")
    # the same domains, one context each, one after the other
    seq_steps = max(1, args.steps // 4)
    t_seq = 0.0
    for d in doms:
        s = pkg.Solver(d["cfg"].solver_params(), d["cfg"].initial_mu, cells_per_thread=k, precision=precision)
        s.set_state(d["phi"], d["c"], d["mu"], [0.0, d["cfg"].dt, 0, 0, 0, 0, 0, 0, 0])
        s.run(1e-6, args.warmup)
        t0 = time.perf_counter()
        s.run(1e-6, seq_steps)
        t_seq += time.perf_counter() - t0
        s.close()
    sequential = cells * seq_steps / t_seq / 1e6
    log(f"sequential: {seq_steps} steps per domain in {t_seq:.3f} s -> {sequential:.1f} Mcell-steps/s")
    out = {
        "metric": "Mcell-steps/s", "value": round(batched, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32", "data": "synthetic",
        "config": {"workload": f"batched ensemble of {len(doms)} uneven domains (seed {args.seed}), "
                               f"{'FP64' if args.precision == 'fp64' else 'FP32-mixed'}, "
                               f"one context, rounds replayed from a captured graph",
                   "domains": len(doms), "cells": cells, "cells_per_thread": k},
        "autotune_Mcell_steps_per_s": {str(kk): round(v, 1) for kk, v in tuned.items()} if tuned else None,
        "sequential_same_domains": round(sequential, 3), "batched_over_sequential": round(batched / sequential, 2),
        "attempts": ctr["attempts"], "cg_iterations": ctr["cg_iterations"],
    }
    print(json.dumps(out), flush=True)


def run_smoke(args, pkg):
    """Config 0 end to end: the reference's own CPU case (256x128, 4 equal
    lamella pairs of 8 cementite + 56 ferrite columns in rows 0-15, 10 K, to
    0.015 us, adaptive dt), run by the drop-in program pearl_gpu and by the
    unmodified reference programs (compiled from source under oracle/_ref)
    on the same input file: pearl_serial on one core and pearl_parallel under
    mpiexec (16 ranks, 4x4).  Each program's step loop is timed from its own
    diagnostics.csv (wall_s of the last output row); both include the
    output-time frames.  Reports the step, rejection and PCG counts of both
    and the final-state differences."""
    import importlib
    import subprocess
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ref_tool
    inputs = importlib.import_module("pearl_amd.inputs")
    tmp = tempfile.mkdtemp()
    inp = os.path.join(tmp, "smoke.in")
    inputs.colony_input(inp, 256, 128, 4, 16, 1.5e-8, output_dt=1.5e-9)
    exe = os.path.join(PKG_DIR, "pearl_gpu")
    ref_dir = os.path.join(ROOT, "oracle", "_ref")

    def run(cmd, out, timeout):
        t0 = time.perf_counter()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"{cmd[0]} failed: {r.stderr.strip()[-500:]}")
        rows = np.genfromtxt(os.path.join(out, "diagnostics.csv"), delimiter=",", names=True)
        status = dict(l.split("=", 1) for l in open(os.path.join(out, "COMPLETED.txt")).read().splitlines() if "=" in l)
        steps = int(rows["step"][-1])
        loop = float(rows["wall_s"][-1])
        log(f"  {os.path.basename(cmd[0])}: {steps} steps, loop {loop:.2f} s, whole run {time.perf_counter() - t0:.2f} s")
        return dict(steps=steps, loop_s=loop, rejected=int(status["rejected"]), cg_total=int(status["CG_total"]),
                    value=256 * 128 * steps / loop / 1e6)

    log("smoke (config 0): drop-in program on the GPU")
    gpu = run([exe, inp, os.path.join(tmp, "gpu")], os.path.join(tmp, "gpu"), 600)
    out = {"metric": "Mcell-steps/s", "value": round(gpu["value"], 3), "unit": "Mcell-steps/s", "n_gpus": 1,
           "steps": gpu["steps"], "warmup": 0, "ms_per_step": round(gpu["loop_s"] * 1e3 / gpu["steps"], 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "config 0: 256x128, 4 equal lamella pairs (8 cementite + 56 ferrite columns, "
                                  "rows 0-15), 10 K, to 0.015 us, adaptive dt; the drop-in program pearl_gpu end to "
                                  "end, step loop from its diagnostics.csv (frames included)",
                      "nx": 256, "ny": 128},
           "gpu": gpu}
    serial = os.path.join(ref_dir, "pearl_serial")
    if os.path.exists(serial) and not args.no_cpu_baseline:
        log("smoke: reference pearl_serial")
        ref = run([serial, inp, os.path.join(tmp, "ref")], os.path.join(tmp, "ref"), 1200)
        out["cpu_baseline_serial"] = dict(value=ref["value"], unit="Mcell-steps/s", cores=1, kind="reference",
                                          sample=f"the whole config-0 run, {ref['steps']} steps, loop "
                                                 f"{ref['loop_s']:.1f} s", **{k: ref[k] for k in ("steps", "rejected",
                                                                                                   "cg_total")})
        a = ref_tool.read_checkpoint(inp, os.path.join(tmp, "ref", "checkpoint.h5"), tmp)
        b = ref_tool.read_checkpoint(inp, os.path.join(tmp, "gpu", "checkpoint.h5"), tmp)
        out["parity"] = {"same_step_rejection_cg_counts": (ref["steps"], ref["rejected"], ref["cg_total"]) ==
                         (gpu["steps"], gpu["rejected"], gpu["cg_total"]),
                         "max_abs_dmu_J_per_mol": float(np.max(np.abs(a[2] - b[2]))),
                         "max_abs_dphi": float(np.max(np.abs(a[0] - b[0])))}
        par, mpi = ref_tool.parallel_program(), ref_tool.mpiexec()
        if par and mpi:
            px, py = ref_tool.grid_split(16, 256, 128)
            log("smoke: reference pearl_parallel, 16 ranks")
            m = run([mpi, "-n", "16", par, inp, os.path.join(tmp, "mpi"), "--px", str(px), "--py", str(py)],
                    os.path.join(tmp, "mpi"), 1200)
            out["cpu_baseline"] = dict(value=m["value"], unit="Mcell-steps/s", cores=16, kind="reference",
                                       sample=f"the whole config-0 run under mpiexec -n 16 ({px}x{py}), "
                                              f"{m['steps']} steps, loop {m['loop_s']:.1f} s")
        else:
            out["cpu_baseline"] = out["cpu_baseline_serial"]
    print(json.dumps(out), flush=True)


def run_adjoint(args, pkg):
    """The checkpointed adjoint on the parameter-recovery case: 512x256, 4
    lamella pairs, 15 K, a frozen dt sequence of --adj-steps steps,
    checkpoints every --adj-every steps (torch_adjoint.colony_inverse_problem).
    Times one pearl_adjoint call (forward with checkpoints, loss, reverse
    sweep with segment recomputation) and the forward alone, and checks the
    gradient against an FP64 autograd reference (the step restated in torch
    with a stencil PCG solve differentiated implicitly, on the same GPU) and
    against central differences of the forward at h = 1e-6 and 1e-8
    relative."""
    import importlib
    import numpy as np
    import torch
    TA = importlib.import_module("pearl_amd.torch_adjoint")
    nx, ny = (512, 256) if (args.nx, args.ny) == (4096, 2048) else (args.nx, args.ny)
    pairs = 4 if args.pairs == 32 else args.pairs
    uc = 15.0 if args.undercooling == 10.0 else args.undercooling
    steps, every = args.adj_steps, args.adj_every
    log(f"adjoint: {nx}x{ny}, {pairs} pairs, {uc} K, {steps} steps, checkpoints every {every}")
    prob, factors, cfg = TA.colony_inverse_problem(nx, ny, pairs, uc, steps, every, seed=args.seed)
    theta = prob.theta0().tolist()
    prob.loss_and_grad(theta)  # warm-up
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        th = prob.theta0().requires_grad_(True)
        L = TA.phase_field_loss(th, prob)
        L.backward()
    wall = (time.perf_counter() - t0) / reps
    grad = th.grad.numpy().copy()
    t0 = time.perf_counter()
    L_fwd = prob.loss(theta)
    wall_fwd = time.perf_counter() - t0
    log(f"  adjoint {wall:.3f} s per call, forward alone {wall_fwd:.3f} s; loss {L.item():.12e}")
    fd = {}
    for rh in (1e-6, 1e-8):
        row = []
        for k, name in enumerate(TA.THETA):
            h = rh * theta[k]
            tp, tm = list(theta), list(theta)
            tp[k] += h
            tm[k] -= h
            d = (prob.loss(tp) - prob.loss(tm)) / (2 * h)
            row.append({"fd": d, "rel": abs(d - grad[k]) / abs(grad[k])})
        fd[f"{rh:.0e}"] = row
        log(f"  central differences h={rh:.0e}: rel " + " ".join(f"{r['rel']:.1e}" for r in row))
    ref = None
    if not args.adj_no_reference:
        import sys as _s
        _s.path.insert(0, os.path.join(ROOT, "tests"))
        import adjoint_reference as R
        phi, c, mu, _ = prob.state
        t0 = time.perf_counter()
        Lr, gr, phr, *_ = R.loss_and_grad(prob.params, prob.mu_reference, phi, c, mu, prob.dts, prob.phi_target,
                                          prob.w_front, prob.front_target, solve="cg", device="cuda", every=every)
        t_ref = time.perf_counter() - t0
        dphi = float(np.max(np.abs(prob.final_phases(theta) - phr)))
        ref = {"seconds": round(t_ref, 2), "loss": Lr, "gradient": [float(v) for v in gr],
               "rel": [float(abs(grad[k] - gr[k]) / abs(gr[k])) for k in range(len(TA.THETA))],
               "final_phases_max_diff": dphi,
               "what": "the step restated in torch (FP64, stencil PCG solve differentiated implicitly, "
                       f"checkpointed every {every} steps) on the same GPU, autograd"}
        log(f"  torch reference {t_ref:.1f} s; rel " + " ".join(f"{v:.1e}" for v in ref["rel"]) +
            f"; final phases differ by {dphi:.1e}")
    out = {
        "metric": "Mcell-steps/s", "value": round(nx * ny * steps / wall / 1e6, 3), "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": round(wall * 1e3 / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"checkpointed adjoint (forward + reverse) of {steps} fixed steps on {nx}x{ny}, "
                               f"{pairs} lamella pairs, {uc} K, checkpoints every {every}; target from the "
                               f"ferrite/austenite and cementite/austenite interface energies and mobility "
                               f"coefficients scaled by seeded factors in [0.8, 1.2]",
                   "nx": nx, "ny": ny, "every": every},
        "factors": factors, "loss": L.item(), "loss_forward_only": L_fwd,
        "theta": dict(zip(TA.THETA, theta)), "gradient": dict(zip(TA.THETA, [float(v) for v in grad])),
        "forward_only_Mcell_steps_per_s": round(nx * ny * steps / wall_fwd / 1e6, 3),
        "adjoint_over_forward_time": round(wall / wall_fwd, 2),
        "central_differences": fd, "torch_reference": ref,
    }
    print(json.dumps(out), flush=True)


def host_loop(s, cfg, args):
    """The host program's loop on solver `s`: steps(count) takes `count`
    accepted steps, pearl_run between output times with the reference loop's
    output tolerance, and the moving-window check at each output time.
    Returns (steps, [shifts, window checks])."""
    next_output = [(math.floor((0.0 + 1e-12 * cfg.output_dt) / cfg.output_dt) + 1.0) * cfg.output_dt]
    shifts = [0, 0]  # shifts, window checks
    eps_out = 1e-12 * max(cfg.end_time, cfg.output_dt)
    s.set_stop_tolerance(eps_out)

    def at_output():
        t = s.scalars().time
        if t >= next_output[0] - eps_out or t >= cfg.end_time - eps_out:
            # the output-time clock snap is applied by the host driver; here
            # only the moving-window check of the hot path is exercised
            sh, _ = s.maybe_shift(cfg.moving_window, cfg.shift_trigger, cfg.shift_target, cfg.shift_gamma_carbon)
            shifts[0] += int(sh)
            shifts[1] += 1
            next_output[0] += cfg.output_dt
            return True
        return False

    def steps(count):
        left = count
        while left > 0:
            if args.host_loop == "advance":
                s.advance(min(cfg.end_time, next_output[0]))
                taken = 1
            else:
                taken = s.run(min(cfg.end_time, next_output[0]), left)["taken"]
            left -= taken
            if not at_output() and taken == 0:
                raise RuntimeError("the run made no progress")

    return steps, shifts


def lockstep_errors(pkg, cfg, phi, c, mu, sc0, n_lock, every, k):
    """FP64 and FP32-mixed contexts from the same state in lockstep: each
    step the FP64 context advances and the mixed context takes a step of
    exactly the accepted dt, so both follow one dt sequence and the
    difference is precision alone.  Every `every` steps (and at the end):
    max |dmu| in J/mol, max |dphi| per phase fraction, max |dc|."""
    import numpy as np
    rows, broken = [], None
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=k) as a, \
            pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=k, precision=pkg.MIXED) as b:
        for z in (a, b):
            z.set_state(phi, c, mu, sc0)
        for step in range(1, n_lock + 1):
            info = a.advance(1.0)
            sc = b.scalars()
            sc.dt = info.accepted_dt
            b.set_scalars(sc)
            r0 = b.counters()["rejected_attempts"]
            b.advance(1.0)
            if broken is None and (b.counters()["rejected_attempts"] != r0 or b.scalars().time != a.scalars().time):
                broken = step
            if step % every == 0 or step == n_lock:
                pa, ca, ma, _ = a.get_state()
                pb, cb, mb, _ = b.get_state()
                rows.append({"step": step, "time": a.scalars().time,
                             "max_abs_dmu_J_per_mol": float(np.max(np.abs(ma - mb))),
                             "max_abs_dphi": [float(np.max(np.abs(pa[i] - pb[i]))) for i in range(3)],
                             "max_abs_dc": float(np.max(np.abs(ca - cb)))})
                log(f"  lockstep step {step}: max|dmu| {rows[-1]['max_abs_dmu_J_per_mol']:.3e} J/mol, "
                    f"max|dphi| {rows[-1]['max_abs_dphi']}")
    return {"lockstep_steps": n_lock, "lockstep_broken_at": broken, "every": rows,
            "max_abs_dmu_J_per_mol": rows[-1]["max_abs_dmu_J_per_mol"], "max_abs_dphi": rows[-1]["max_abs_dphi"]}


def mixed_beside(pkg, cfg, phi, c, mu, initial_carbon, args, have_torch, torch, nx, ny, tmp):
    """The FP32-mixed form beside the FP64 one: its throughput on the same
    workload over the same --steps accepted steps (same host loop), and its
    error against the FP64 form over --mixed-lockstep steps in lockstep
    (independent of --steps), on this workload and on the 8192x4096 one
    (64 pairs, 20 K undercooling, same seed), reported every 1,000 steps."""
    sc0 = [0.0, cfg.dt, 0.0, initial_carbon, 0.0, 0.0, 0, 0, 0]
    out = {}
    with pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=args.cells_per_thread,
                    precision=pkg.MIXED) as m:
        m.set_state(phi, c, mu, sc0)
        steps, _ = host_loop(m, cfg, args)
        steps(args.warmup)
        c0 = m.counters()
        if have_torch:
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        steps(args.steps)
        if have_torch:
            torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        c1 = m.counters()
        out.update(value=round(nx * ny * args.steps / wall / 1e6, 3), ms_per_step=round(wall * 1e3 / args.steps, 3),
                   steps=args.steps, attempts=c1["attempts"] - c0["attempts"],
                   cg_iterations=c1["cg_iterations"] - c0["cg_iterations"])
    n_lock = args.mixed_lockstep
    if n_lock > 0:
        log(f"mixed form: lockstep against FP64 on {nx}x{ny}, {n_lock} steps")
        out["lockstep"] = lockstep_errors(pkg, cfg, phi, c, mu, sc0, n_lock, 1000, args.cells_per_thread)
        out["lockstep"]["workload"] = f"{nx}x{ny}, {args.pairs} pairs, {args.undercooling} K (this workload)"
        nx2, ny2, pairs2, uc2 = 8192, 4096, 64, 20.0
        log(f"mixed form: lockstep against FP64 on {nx2}x{ny2}, {pairs2} pairs, {uc2} K, {n_lock} steps")
        # a directory of its own: the workload's input file must not replace this run's
        _, cfg2, phi2, c2, mu2 = build_workload(pkg, nx2, ny2, pairs2, ny2 // 8, args.seed, uc2,
                                                tempfile.mkdtemp(dir=tmp))
        with pkg.Solver(cfg2.solver_params(), cfg2.initial_mu) as z:
            z.set_state(phi2, c2, mu2, [0.0, cfg2.dt, 0.0, 0.0, 0.0, 0.0, 0, 0, 0])
            m2 = z.total_carbon()
        lk = lockstep_errors(pkg, cfg2, phi2, c2, mu2, [0.0, cfg2.dt, 0.0, m2, 0.0, 0.0, 0, 0, 0], n_lock, 1000,
                             args.cells_per_thread)
        lk["workload"] = f"{nx2}x{ny2}, {pairs2} pairs +/-5% (seed {args.seed}), {uc2} K"
        out["lockstep_8192x4096_20K"] = lk
        out["max_abs_dmu_J_per_mol"] = max(out["lockstep"]["max_abs_dmu_J_per_mol"], lk["max_abs_dmu_J_per_mol"])
        del phi2, c2, mu2
    log(f"mixed form: {out['value']:.1f} Mcell-steps/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000,
                    help="accepted steps in the timed region (the production workload is 10,000)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--undercooling", type=float, default=10.0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="production: skip the FP32-mixed run beside the FP64 one")
    ap.add_argument("--mixed-lockstep", type=int, default=1000,
                    help="production: lockstep steps for the mixed form's stated error (on this workload and on "
                         "the 8192x4096, 64-pair, 20 K one; independent of --steps; 0 skips it)")
    ap.add_argument("--kernel-steps", type=int, default=500,
                    help="accepted steps of the per-kernel event region after the timed one")
    ap.add_argument("--no-profile", action="store_true",
                    help="skip the second region with per-kernel events (no roofline)")
    ap.add_argument("--workload", choices=["production", "ensemble", "mixed-accuracy", "adjoint", "smoke"],
                    default="production")
    ap.add_argument("--shift-trigger", type=float, default=0.1,
                    help="production: moving-window trigger as a fraction of the domain height (default 0.1, "
                         "below the colony front at 1/8 of the height, so the window fires inside the timed "
                         "region; the input default 0.65 is not reached in 10,000 steps; 0 keeps it)")
    ap.add_argument("--shift-target", type=float, default=None,
                    help="production: shift target as a fraction of the height (default: half the trigger)")
    ap.add_argument("--output-dt", type=float, default=None,
                    help="production: output interval in seconds, at which the window check runs (default 3e-10, "
                         "about every 100 steps)")
    ap.add_argument("--host-loop", choices=["run", "advance"], default="run",
                    help="production: pearl_run between output times (the host program's loop) or one "
                         "pearl_advance call per step")
    ap.add_argument("--adj-steps", type=int, default=2000, help="adjoint: fixed steps")
    ap.add_argument("--adj-every", type=int, default=100, help="adjoint: checkpoint interval")
    ap.add_argument("--adj-no-reference", action="store_true", help="adjoint: skip the torch autograd reference")
    ap.add_argument("--precision", choices=["fp64", "mixed", "stencil32"], default="fp64",
                    help="mixed: Jacobi diagonal and PCG vectors p, z stored in FP32 (state, operator "
                         "coefficients, residuals, dots, mass and acceptance FP64)")
    ap.add_argument("--acc-steps", type=int, default=3000, help="mixed-accuracy: FP64 steps in total")
    ap.add_argument("--acc-every", type=int, default=1000, help="mixed-accuracy: steps between comparisons")
    ap.add_argument("--acc-mode", choices=["lockstep", "free"], default="lockstep",
                    help="mixed-accuracy: one dt sequence, or each form on its own dt compared at matched times")
    ap.add_argument("--acc-other", choices=["mixed", "fp64-k8", "stencil32"], default="mixed",
                    help="mixed-accuracy: the form compared with FP64 at the default tile height; fp64-k8 is the "
                         "FP64 form with another summation order (8-row tiles), the noise floor of the comparison")
    ap.add_argument("--domains", type=int, default=64)
    ap.add_argument("--cells-per-thread", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 1 GiB device copy (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs one domain on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    pkg = load_pkg()
    if args.workload == "ensemble":
        run_ensemble(args, pkg)
        return
    if args.workload == "mixed-accuracy":
        run_mixed_accuracy(args, pkg)
        return
    if args.workload == "adjoint":
        run_adjoint(args, pkg)
        return
    if args.workload == "smoke":
        run_smoke(args, pkg)
        return
    try:
        import torch
        have_torch = torch.cuda.is_available()
    except Exception:  # noqa: BLE001
        torch, have_torch = None, False

    nx, ny = args.nx, args.ny
    strip_rows = ny // 8
    tmp = tempfile.mkdtemp()
    window = None
    window = {}
    if args.shift_trigger:
        # a lower trigger than the input default (0.65 of the height), so the
        # front detection, strip check and row shift all run at output times
        import importlib
        dy = float(dict(importlib.import_module("pearl_amd.inputs").DEFAULT_PARAMETERS)["dx"])
        tgt = args.shift_target if args.shift_target is not None else args.shift_trigger / 2
        window = {"shift_trigger": repr(args.shift_trigger * ny * dy), "shift_target": repr(tgt * ny * dy)}
    if args.output_dt:
        window["output_dt"] = repr(args.output_dt)
    path, cfg, phi, c, mu = build_workload(pkg, nx, ny, args.pairs, strip_rows, args.seed, args.undercooling, tmp,
                                           overrides=window)
    log(f"workload {nx}x{ny}: {args.pairs} pairs, colony {strip_rows} rows, {args.undercooling} K")
    precision = {"fp64": pkg.FP64, "mixed": pkg.MIXED, "stencil32": pkg.FP32_STENCIL}[args.precision]
    s = pkg.Solver(cfg.solver_params(), cfg.initial_mu, cells_per_thread=args.cells_per_thread, precision=precision)
    # The initial state is formed on the device: the colony's separable
    # profiles (initialize's lamella branch) and the carbon at the common mu
    # as initialize forms it; the host copies feed the other legs.
    initial = importlib.import_module("pearl_amd.initial")
    th, prod = initial.colony_profiles(nx, ny, cfg.dx, cfg.dy, cfg.eps, initial.plates_from_fills(cfg.filling),
                                       strip_rows)
    s.init_profile(pkg.PROFILE_LAMELLA, th, prod)
    _, initial_carbon, valid = s.init_composition(pkg.COMPOSITION_MU, cfg.initial_mu)
    if not valid:
        raise RuntimeError("initial state out of model range")
    s.set_scalars([0.0, cfg.dt, 0.0, initial_carbon, 0.0, 0.0, 0, 0, 0])
    phi, c, mu, _ = s.get_state()

    steps, shifts = host_loop(s, cfg, args)

    if args.copy_calibration and have_torch:
        # A plain device-to-device copy of 1 GiB: FETCH_SIZE / WRITE_SIZE of
        # this kernel scale the counters of the solver kernels.
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    steps(args.warmup)
    log(f"warmup {args.warmup} steps done, t = {s.scalars().time:.4e}")
    # Timed region: exactly --steps accepted steps, no per-kernel events
    # (recording events around every kernel of a round adds graph nodes and
    # slows the loop by ~15 %).
    c0 = s.counters()
    s.profile(False)
    if have_torch:
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps(args.steps)
    if have_torch:
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c1 = s.counters()
    counted = s.kernel_stats()  # launches and algorithmic bytes of the timed region
    sc = s.scalars()
    d = s.diagnostics()
    log(f"timed {args.steps} steps in {wall:.3f} s; t = {sc.time:.4e}, rejected attempts "
        f"{c1['rejected_attempts'] - c0['rejected_attempts']}, CG {c1['cg_iterations'] - c0['cg_iterations']}")
    # Kernel region: the next --kernel-steps steps with HIP events around each
    # cell kernel on the library's stream; the per-kernel averages for the
    # roofline.
    stats, wall_ev = {}, None
    if not args.no_profile:
        s.profile(True)
        if have_torch:
            torch.cuda.synchronize()
        t1 = time.perf_counter()
        steps(args.kernel_steps)
        if have_torch:
            torch.cuda.synchronize()
        wall_ev = time.perf_counter() - t1
        stats = s.kernel_stats()
        s.profile(False)
        log(f"kernel region: {args.kernel_steps} more steps with per-kernel events in {wall_ev:.3f} s")
        for name, v in stats.items():
            if v["launches"]:
                log(f"  {name:16s} {v['launches']:6d} launches {v['ms']:9.3f} ms  "
                    f"{v['bytes'] / max(v['ms'], 1e-9) / 1e6:8.1f} GB/s")
    # Measured bandwidth ceiling beside the vendor peak: a hand-written plain
    # copy of 1 GiB (pearl_stream_probe; read + write counted).
    copy_gbs, mix = None, {}
    if not args.no_profile:
        copy_gbs = pkg.stream_probe(0, 1 << 30, 20)
        log(f"copy ceiling: {copy_gbs:.1f} GB/s (1 GiB device copy, pearl_stream_probe)")
        if args.precision == "fp64":
            # each kernel's own access pattern without its arithmetic (for
            # the kernels that moved bytes in the region)
            for name, (r, w) in pkg.KERNEL_STREAMS.items():
                if not stats.get(name, {}).get("bytes"):
                    continue
                mix[name] = pkg.stream_mix(r, w, nx, ny, s.layout()[1])
                log(f"  bare pattern {name:16s} {r} read + {w} written planes: {mix[name][0]:7.1f} us, "
                    f"{mix[name][1]:7.1f} GB/s")
    mcs = nx * ny * args.steps / wall / 1e6
    rows_per_tile = s.layout()[1]
    roof = None
    active = {k: v for k, v in stats.items() if v["ms"] > 0}
    if active:
        dom = max(active.items(), key=lambda kv: kv[1]["ms"])
        achieved = dom[1]["bytes"] / (dom[1]["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "copy_ceiling_GBps": round(copy_gbs, 1) if copy_gbs else None,
                "frac_of_copy": round(achieved / copy_gbs, 4) if copy_gbs else None,
                "bytes_per_launch": dom[1]["bytes"] / dom[1]["launches"],
                "avg_launch_ms": dom[1]["ms"] / dom[1]["launches"],
                "region": f"the {args.kernel_steps} steps after the timed region, HIP events around each cell kernel",
                "kernels": {k: {"achieved_GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1),
                                "avg_launch_ms": round(v["ms"] / v["launches"], 5), "launches": v["launches"]}
                            for k, v in active.items() if v["bytes"] > 0}}
        for k, v in roof["kernels"].items():
            if k in mix:
                # a bare kernel with this kernel's planes, no arithmetic: the
                # rate its access pattern reaches (a kernel with better issue,
                # e.g. two threads per column, can beat it)
                v["bare_pattern_GBps"] = round(mix[k][1], 1)
                v["bare_pattern_over_kernel_time"] = round(mix[k][0] / (v["avg_launch_ms"] * 1e3), 4)
        measured = os.path.join(PROFILES, f"hbm_traffic_{args.precision}_{nx}x{ny}_k{rows_per_tile}.json")
        if os.path.exists(measured):
            with open(measured) as fh:
                tr = json.load(fh)["kernels"].get(dom[0])
            if tr:
                # the counter run's traffic per algorithmic byte, applied to
                # this run's launches (a kernel with several launch types has
                # another mix of them here)
                ratio = tr["traffic_bytes"] / tr["algorithmic_bytes"]
                roof["traffic"] = round(ratio * roof["bytes_per_launch"])
                roof["traffic_over_algorithmic"] = round(ratio, 4)
                roof["traffic_source"] = os.path.relpath(measured, ROOT)
                # measured HBM rate (counter bytes per launch / launch time)
                roof["traffic_GBps"] = round(roof["traffic"] / (roof["avg_launch_ms"] * 1e-3) / 1e9, 1)
                roof["traffic_frac"] = round(roof["traffic_GBps"] / HBM_PEAK_GBS, 4)
    total_bytes = sum(v["bytes"] for v in counted.values())
    out = {
        "metric": "Mcell-steps/s", "value": round(mcs, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64+f32", "data": "synthetic",
        "config": {"workload": f"{'FP64 plain' if args.precision == 'fp64' else 'FP32-mixed'} form, {nx}x{ny} "
                               f"lamellar colony ({args.pairs} pairs, +/-5% spacing, "
                               f"seed {args.seed}), {args.undercooling} K undercooling, adaptive dt, implicit "
                               f"diffusion, moving window",
                   "nx": nx, "ny": ny, "precision": args.precision, "cells_per_thread": rows_per_tile,
                   "host_loop": args.host_loop, "shift_trigger": cfg.shift_trigger,
                   "shift_target": cfg.shift_target, "output_dt": cfg.output_dt},
        "attempts": c1["attempts"] - c0["attempts"],
        "rejected_attempts": c1["rejected_attempts"] - c0["rejected_attempts"],
        "cg_iterations": c1["cg_iterations"] - c0["cg_iterations"],
        "loop_algorithmic_GBps": round(total_bytes / wall / 1e9, 1),
        "with_kernel_events_Mcell_steps_per_s": round(nx * ny * args.kernel_steps / wall_ev / 1e6, 3) if wall_ev else None,
        "roofline": roof,
        "state_check": {"time": sc.time, "step": sc.step, "phi_sum_error": d["phi_sum_error"],
                        "mass_ledger": d["mass"] + sc.carbon_removed - sc.carbon_added - sc.initial_carbon,
                        "shifts": shifts[0], "window_checks": shifts[1]},
    }
    s.close()
    if args.precision == "fp64" and not args.no_mixed:
        # the metric's other half: the FP32-mixed form at its stated mu error
        mx = mixed_beside(pkg, cfg, phi, c, mu, initial_carbon, args, have_torch, torch if have_torch else None, nx, ny,
                          tmp)
        mx["over_fp64"] = round(mx["value"] / out["value"], 3)
        out["mixed"] = mx
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"], out["cpu_baseline_serial"], out["parity_vs_reference"] = cpu_baseline(
            path, cfg, phi, c, mu, initial_carbon, args.cpu_seconds, tmp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
