#!/usr/bin/env python3
"""Benchmark of the plain WCSPH step on one MI355X.

Workload: the 1,048,576-particle dam break (128x64x128 lattice, spacing
0.01, h = 0.02, 2x1x2 box, velocities N(0, 0.05), dt = 5e-5).  A step is
one full pass of the hot path: integrate + hash, the new (cell, index)
order with cellStart/cellEnd (built from the last order; the stable radix
sort when a cell crowds), reorder, density + EOS, force + velocity update.  The
state is resident in HBM before the timed region; nothing is copied to or
from the host inside it.

Prints ONE JSON line (rank 0, N = 1):
  value        particle-updates/s = N_particles * steps / timed seconds, in
               the precision mode --prec names (default fast: integer results
               exact, floats within 1e-5 of the reference arithmetic and of
               FP64; "reference" is bit-exact with the FP32 restatement)
  modes        the same workload in both modes (fast and reference order):
               step rate and per-kernel device times
  roofline     dominant gather kernel: algorithmic bytes per launch / mean
               launch time (HIP events on the context stream) against 8 TB/s
  cpu_baseline the CPU oracle (FP32 reference order) on the first steps of the
               same 1M workload, all host threads, bounded in time (the
               single-thread rate on the 15,625-particle dam break over its
               100 steps rides along)
  learned      the learned surrogate: trained on the device from plain-solver
               trajectories, then stepping 1,000 dt of simulated time at 4 x dt
               from the same state as a 1,000-step plain run (whatever --steps
               is): simulated-time throughput against that plain run, density
               error against it, beside the plain step at 4 x dt and the zero
               network, and the surrogate kernel's matrix-core roofline (BF16
               dense peak)
  batch        64 uneven scenes advanced as one array by an autotuned,
               graph-captured step (particle-updates/s over all scenes)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from pibiti_amd import scenes, sph  # noqa: E402

METRIC = "particle-updates/s at 1M (plain vs learned, ≤1% ρ err); % HBM peak"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense


def progress(msg: str) -> None:
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def pair_flops(pairs: float, n: int) -> float:
    """Algorithmic FLOPs of one surrogate pass: both pair layers for every
    listed pair over the 9 real edge inputs (9 -> 64 -> 64, 2 FLOPs per
    multiply-add) and the node layer (67 -> 3) per particle.  The matrix
    cores run the first layer 16 inputs wide (the padding is not counted)."""
    return 2.0 * pairs * (9 * 64 + 64 * 64) + 2.0 * n * 67 * 3


def learned_section(args, p, snap_pos, snap_vel) -> dict:
    """Train the surrogate, then advance 1,000 dt of simulated time from the
    state after the warm-up four ways: the plain step at dt (the reference
    trajectory and the throughput baseline), the surrogate at 4 x dt with the
    1 % fallback, the plain step at 4 x dt and the zero network."""
    from pibiti_amd import learn
    n = int(p.num_particles)
    rho0 = float(p.rest_density)
    factor = 4
    fine = args.learned_dt
    coarse = fine // factor
    progress(f"learned: training ({args.train_steps} steps, {args.train_iters} updates per window)")
    model = sph.Model(learn.default_norm(p), seed=1)
    hist = learn.train(model, scenes_count=200, steps=args.train_steps, stride=8, iters=args.train_iters,
                       lr=args.train_lr, lr_end=1e-4)
    L = hist["loss"]
    tail = L[-max(1, len(L) // 10):]

    q = p.copy()
    q.tune[2] = args.check  # surrogate steps per host check and per list build

    def fresh(params=q):
        c = sph.SphSystem(params)
        c.reset(snap_pos, snap_vel)
        c.update(1)  # the state the timed plain run started from
        return c

    def err(rho, ref):
        d = np.abs(np.asarray(rho, np.float64) - ref) / rho0
        return {"rms": float(np.sqrt(np.mean(d * d))), "max": float(d.max())}

    # the plain run at dt: reference trajectory and throughput baseline
    progress(f"learned: plain run over {fine} dt")
    with fresh() as c:
        c.synchronize()
        t0 = time.perf_counter()
        c.update(fine)
        c.synchronize()
        plain_secs = time.perf_counter() - t0
        rho_plain_end = c.get_array(sph.SPH_ARRAY_DENSITY).astype(np.float64)
    # one surrogate step, then one plain step (so that the network's velocity
    # update has moved the particles), against factor + 1 plain steps
    with fresh() as c:
        c.update(factor + 1)
        rho_ref1 = c.get_array(sph.SPH_ARRAY_DENSITY).astype(np.float64)
    with fresh() as c:
        c.step_learned(model, 1, factor=factor, tol=1e9, skin=args.skin)
        c.update(1)
        one = err(c.get_array(sph.SPH_ARRAY_DENSITY), rho_ref1)
    # the whole workload, with the 1 % drift fallback
    progress(f"learned: surrogate run, {coarse} steps of {factor} dt")
    with fresh() as c:
        # zero steps: allocates the surrogate's buffers outside the timed region
        c.step_learned(model, 0, factor=factor, tol=args.tol, skin=args.skin)
        c.synchronize()
        t0 = time.perf_counter()
        rep = c.step_learned(model, coarse, factor=factor, tol=args.tol, skin=args.skin)
        c.synchronize()
        secs = time.perf_counter() - t0
        run = err(c.get_array(sph.SPH_ARRAY_DENSITY), rho_plain_end)
    # the plain step at factor x dt over the same workload
    progress("learned: plain step at 4 dt, zero network, kernel timings")
    big = q.copy()
    big.time_step = factor * float(p.time_step)
    with fresh() as c:
        c.set_params(big)
        c.synchronize()
        t0 = time.perf_counter()
        c.update(coarse)
        c.synchronize()
        big_secs = time.perf_counter() - t0
        run_plain_big = err(c.get_array(sph.SPH_ARRAY_DENSITY), rho_plain_end)
    # the zero network: the coarse step without any pair force
    zero = sph.Model(learn.default_norm(p), seed=1)
    zero.set_weights(np.zeros(sph.MLP_NUM_PARAMS, np.float32))
    with fresh() as c:
        c.step_learned(zero, coarse, factor=factor, tol=1e9, skin=args.skin)
        run_zero = err(c.get_array(sph.SPH_ARRAY_DENSITY), rho_plain_end)
    with fresh() as c:
        c.step_learned(zero, 1, factor=factor, tol=1e9, skin=args.skin)
        c.update(1)
        one_zero = err(c.get_array(sph.SPH_ARRAY_DENSITY), rho_ref1)
    zero.close()
    # kernel times (HIP events) over a few more surrogate steps
    with fresh() as c:
        c.set_timing(True)
        krep = c.step_learned(model, 40, factor=factor, tol=1e9, skin=args.skin)
    sim_steps = factor * rep["coarse_steps"] + rep["plain_steps"]
    speedup = (sim_steps / secs) / (fine / plain_secs)
    flops = pair_flops(krep["pairs"], n)
    achieved = flops / (krep["mlp_ms"] * 1e-3) / 1e12
    model.close()
    return {
        "factor": factor,
        "workload": f"{fine} dt of simulated time ({coarse} surrogate steps of {factor} dt) from the state "
                    "after the warm-up, against a {fine}-step plain run from the same state".replace("{fine}",
                                                                                                 str(fine)),
        "value": n * sim_steps / secs,
        "unit": "particle-updates/s (dt-equivalent)",
        "sim_time_speedup": round(speedup, 3),
        "plain_ms_per_step": plain_secs * 1e3 / fine,
        "ms_per_coarse_step": secs * 1e3 / max(1, rep["coarse_steps"] + rep["fallbacks"]),
        "fallbacks": rep["fallbacks"],
        "coarse_steps": rep["coarse_steps"],
        "tol": args.tol,
        "skin": args.skin,
        "steps_per_list_build": args.check if args.skin > 0 else 1,
        "list_builds": rep["rebuilds"],
        "early_list_builds": rep["early_rebuilds"],
        "density_error_one_step": one,
        "density_error_workload": run,
        "density_error_note": "|rho - rho_plain| / rho0 over particles: one surrogate step and one plain "
                              "step after it against 5 plain steps from the same state (so that the network's "
                              "velocity update has moved the particles), and the whole workload against the "
                              "plain run",
        "plain_at_factor_dt": {"density_error_workload": run_plain_big,
                               "sim_time_speedup": round((fine / big_secs) / (fine / plain_secs), 3),
                               "note": "the plain step at 4 x dt over the same workload: the method without "
                                       "the network"},
        "zero_network": {"density_error_workload": run_zero, "density_error_one_step": one_zero,
                         "note": "the surrogate step with every weight zero: no pair force at all"},
        "meets_1pct": run["rms"] <= 0.01,
        "beats_plain_at_factor_dt": run["rms"] < run_plain_big["rms"] and speedup > (fine / big_secs) / (fine / plain_secs),
        "roofline": {"bound": "mfma", "kernel": "surrogate pair network",
                     "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 5), "traffic": measured_traffic().get("surrogate"),
                     "pairs_per_launch": krep["pairs"], "flops_per_launch": flops,
                     "flops_convention": "9 real edge inputs (the matrix cores run 16, padded)",
                     "mean_launch_ms": round(krep["mlp_ms"], 5)},
        "kernels_ms": {"density_lists": round(krep["density_ms"], 5), "surrogate": round(krep["mlp_ms"], 5),
                       "list_build": round(krep["rebuild_ms"], 5), "coarse_step": round(krep["coarse_ms"], 5)},
        "training": {"seconds": round(hist["seconds"], 2), "trajectories": hist["scenes"],
                     "particles": hist["particles"], "plain_steps": hist["steps"],
                     "updates": len(L) * hist["iters"], "final_loss": float(np.mean(tail)),
                     "first_loss": float(L[0])},
    }


def batch_section(prec_name: str) -> dict:
    """The batched workload: 64 independent dam breaks with particle counts
    log-uniform in [2k, 200k] (seeded), boxes scaled to each count, advanced
    as one particle array in the given precision mode.  The autotuner picks
    the launch shape on the live state, 10 steps are captured in a hipGraph,
    and 10 simulated ms per scene (100 steps of 1e-4 s) are timed as 10
    graph replays."""
    counts = scenes.batch_counts(64)
    sc = [scenes.batch_scene(c, i, precision=PREC[prec_name]) for i, c in enumerate(counts)]
    params = [q for q, _, _, _ in sc]
    total = sum(int(q.num_particles) for q in params)
    with sph.SphBatch(params) as b:
        for k, (_, x, v, _) in enumerate(sc):
            b.reset(k, x, v)
        b.update(2)
        tune, tuned_ms = b.autotune(None)
        b.capture(10)
        b.replay(1)
        b.synchronize()
        t0 = time.perf_counter()
        b.replay(10)
        b.synchronize()
        secs = time.perf_counter() - t0
        info = b.info()
    return {
        "workload": "64 dam breaks, particle counts log-uniform in [2k, 200k] (seed 5), spacing 0.02, "
                    "h 0.04, dt 1e-4, 10 simulated ms per scene as 10 replays of a 10-step hipGraph",
        "precision": prec_name, "scenes": len(sc), "particles": total, "cells": info["cells"], "steps": 100,
        "value": total * 100 / secs, "unit": "particle-updates/s", "ms_per_step": secs * 1e3 / 100,
        "tune": tune, "autotuned_ms_per_step": round(tuned_ms, 5),
        "min_scene": min(counts), "max_scene": max(counts),
    }


def traffic_file(prec: str = "reference") -> Path:
    suffix = "" if prec == "reference" else f"_{prec}"
    return ROOT / "profiles" / f"hbm_traffic_dam_break_1m{suffix}.txt"


def measured_traffic(prec: str = "reference") -> dict:
    """HBM bytes per launch of the dominant kernels, from the committed
    counter summary of that precision mode (lines 'traffic <kernel>
    <bytes>'), or {} if absent."""
    out = {}
    f = traffic_file(prec)
    if f.exists():
        for line in f.read_text().splitlines():
            f = line.split()
            if len(f) == 3 and f[0] == "traffic":
                out[f[1]] = int(f[2])
    return out


def copy_calibration() -> None:
    """Five device-to-device copies of 256 MiB (the HIP runtime's copy
    kernel): under a memory-counter profile their FETCH_SIZE / WRITE_SIZE
    read against the known 256 MiB."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    size = C.c_size_t(256 << 20)
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), size) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, size) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(5):
            if hip.hipMemcpy(b, a, size, 3) != 0:  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def algorithmic_bytes(p) -> dict:
    """Minimal FP32 traffic per particle of each stage (pos 12 B, vel 12 B,
    rho 4 B, p 4 B, uint32 key and index; every array counted once per
    kernel; cell tables 8 B per cell).  integrate_hash: state read, last key
    read, new key written (the scatter stores the integrated state); sort:
    the cell scan (old ranges
    and mover counts read, new ranges and cleared counts written, 24 B per
    cell, plus one mover bit per particle); reorder: the scatter (new key,
    index and state read, state, index and key written, both cell tables
    read)."""
    n = int(p.num_particles)
    c = int(p.num_cells) / n
    return {
        "integrate_hash": 12 + 12 + 4 + 4,
        "sort": 24 * c + 0.125,
        "reorder": 8 + 24 + 24 + 8 + 16 * c,
        "density": 12 + 8 + 8 * c,
        "force": 32 + 12 + 8 * c,
    }


def cpu_baseline(budget_s: float, p, pos, vel) -> dict:
    """The CPU oracle (FP32 reference order, OpenMP over all host threads) on
    the first steps of the same 1M-particle workload, bounded by budget_s.
    The single-thread rate on the 15,625-particle dam break rides along."""
    import oracle
    oracle.build()
    n = int(p.num_particles)
    threads = oracle.num_threads()
    t0 = time.perf_counter()
    steps = 0
    x, v = pos, vel
    while steps < 100 and (time.perf_counter() - t0) < budget_s:
        o = oracle.run_f32(p, x, v, 1)
        x, v = o["pos"], o["vel"]
        steps += 1
    dt = time.perf_counter() - t0

    q, qx, qv, _ = scenes.cpu_dam_break_15k()
    oracle.set_threads(1)
    t1 = time.perf_counter()
    oracle.run_f32(q, qx, qv, 100)
    dt1 = time.perf_counter() - t1
    oracle.set_threads(threads)
    return {
        "value": n * steps / dt,
        "unit": "particle-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (FP32 reference order) on the first {steps} steps of the same "
                  f"1,048,576-particle dam break ({dt:.1f} s), {threads} OpenMP threads, "
                  f"host {os.uname().nodename}, {os.cpu_count()} logical CPUs visible",
        "cpu_model": cpu_model(),
        "single_thread_15k": {"value": int(q.num_particles) * 100 / dt1,
                              "sample": f"15,625-particle dam break, its 100 steps ({dt1:.1f} s), 1 thread"},
    }


PREC = {"reference": sph.SPH_PREC_REFERENCE, "kahan": sph.SPH_PREC_KAHAN, "fast": sph.SPH_PREC_FAST}


def plain_run(prec_name: str, args) -> dict:
    """The timed plain workload in one precision mode: warm-up, args.steps
    timed steps, then per-kernel device times over args.roofline_steps
    steps.  Returns the rate, the kernel split and the end state (plus the
    snapshot one step before the timed region)."""
    p, pos, vel, info = scenes.gpu_dam_break_1m(precision=PREC[prec_name])
    n = int(p.num_particles)
    s = sph.SphSystem(p)
    s.reset(pos, vel)
    s.update(max(0, args.warmup - 1))
    snap_pos = s.get_array(sph.SPH_ARRAY_POSITION)
    snap_vel = s.get_array(sph.SPH_ARRAY_VELOCITY)
    s.update(1 if args.warmup > 0 else 0)
    s.synchronize()
    t0 = time.perf_counter()
    s.update(args.steps)
    s.synchronize()
    secs = time.perf_counter() - t0
    # per-kernel device times: HIP events on the context stream, one step per call
    s.set_timing(True)
    acc = {}
    for _ in range(args.roofline_steps):
        s.update(1)
        for k, v in s.timing().items():
            acc[k] = acc.get(k, 0.0) + v
    s.set_timing(False)
    ms = {k: v / args.roofline_steps for k, v in acc.items()}
    st = s.stats()
    s.close()
    return dict(p=p, pos=pos, vel=vel, info=info, n=n, secs=secs, ms=ms, stats=st, snap_pos=snap_pos,
                snap_vel=snap_vel, value=n * args.steps / secs, ms_per_step=secs * 1e3 / args.steps)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--prec", choices=list(PREC), default="fast",
                    help="precision mode of the headline plain step (the other of fast/reference is "
                         "measured beside it)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-steps", type=int, default=50)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-modes", action="store_true", help="skip the second precision mode")
    ap.add_argument("--learned-dt", type=int, default=1000,
                    help="length of the surrogate workload in plain time steps (a multiple of 4)")
    ap.add_argument("--train-steps", type=int, default=1600)
    ap.add_argument("--train-iters", type=int, default=32)
    ap.add_argument("--train-lr", type=float, default=3e-3)
    ap.add_argument("--tol", type=float, default=0.01)
    ap.add_argument("--skin", type=float, default=0.0,
                    help="surrogate steps between sorts walk cells within (1 + skin) h (0 = sort every step)")
    ap.add_argument("--check", type=int, default=4,
                    help="surrogate steps per host check (and per sort with a skin)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with five device copies of 256 MiB (a plain copy kernel of known size, "
                         "for scaling memory counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not a build)", file=sys.stderr)
        return 2

    progress(f"plain step, {args.prec} mode: {args.warmup} warm-up and {args.steps} timed steps")
    head = plain_run(args.prec, args)
    p, n, ms, st, info = head["p"], head["n"], head["ms"], head["stats"], head["info"]
    ab = algorithmic_bytes(p)
    kern_gbs = {k: ab[k] * n / (ms[k] * 1e-3) / 1e9 for k in ab if ms.get(k, 0) > 0}
    dom = max(("density", "force"), key=lambda k: ms[k])
    traffic = measured_traffic(args.prec)
    roof = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(kern_gbs[dom], 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(kern_gbs[dom] / HBM_PEAK_GBS, 5),
        "traffic": traffic.get(dom),
        "algorithmic_bytes_per_particle": round(ab[dom], 3),
        "mean_launch_ms": round(ms[dom], 5),
    }
    gather_ms = ms["density"] + ms["force"]
    gather_bytes = (ab["density"] + ab["force"]) * n
    step_bytes = sum(ab.values()) * n

    def mode_summary(r) -> dict:
        m = r["ms"]
        return {"value": r["value"], "ms_per_step": r["ms_per_step"],
                "kernels_ms": {k: round(v, 5) for k, v in m.items()},
                "gather_ms": round(m["density"] + m["force"], 5),
                "gather_hbm_frac": round(gather_bytes / ((m["density"] + m["force"]) * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}

    modes = {args.prec: mode_summary(head)}
    if not args.no_modes:
        other = "reference" if args.prec != "reference" else "fast"
        progress(f"plain step, {other} mode")
        modes[other] = mode_summary(plain_run(other, args))
    modes["note"] = ("fast: integer results exact, floats within 1e-5 of the reference arithmetic and of FP64 "
                     "(condition-scaled); reference: bit-exact with the FP32 restatement of the reference")

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered lattice, N(0,0.05) velocities)",
        "config": {
            "workload": "WCSPH dam break, 1,048,576 particles (128x64x128, spacing 0.01, h 0.02, "
                        "2x1x2 box, dt 5e-5), plain step",
            "particles": n,
            "grid": list(info["grid"]),
            "key_bits": int(p.key_bits),
            "precision": args.prec,
            "eos": "linear",
        },
        "roofline": roof,
        "kernels_ms": {k: round(v, 5) for k, v in ms.items()},
        "kernels_gbs": {k: round(v, 1) for k, v in kern_gbs.items()},
        "gather_hbm_frac": round(gather_bytes / (gather_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "step_hbm_frac": round(step_bytes / (ms["step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "modes": modes,
        "end_state": {"sim_time": st["sim_time"], "max_speed": st["max_speed"],
                      "cap_cells": st["cap_cells"], "max_cell_count": st["max_cell_count"],
                      "cap_truncation": "accepted: cells above max_par_in_cell (16) offer only their first "
                                        "16 slots as candidates, as the reference's walk does "
                                        "(Kernel_Cell.cui:151); counted in cap_cells, reproduced exactly"},
    }
    if not args.no_learned and args.warmup > 0:
        out["learned"] = learned_section(args, p, head["snap_pos"], head["snap_vel"])
    if not args.no_batch:
        progress(f"batch of 64 scenes, {args.prec} mode")
        out["batch"] = batch_section(args.prec)
        other = "reference" if args.prec == "fast" else "fast"
        if not args.no_modes:
            progress(f"batch of 64 scenes, {other} mode")
            out["batch"]["other_mode"] = {k: v for k, v in batch_section(other).items()
                                          if k in ("precision", "value", "ms_per_step", "tune")}
    if not args.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget, p, head["pos"], head["vel"])
    if args.copy_calibration:
        copy_calibration()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
