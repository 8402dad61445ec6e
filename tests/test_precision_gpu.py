"""Precision study at 4,194,304 particles: two colliding blocks at +-1 m/s.

The blocks are advanced on the GPU until they interpenetrate, then one step
is taken in both accumulation modes (reference order and Kahan-compensated
FP32).  Each is checked bit for bit against the FP32 oracle, and its
per-field error against the FP64 oracle on the same post-integrate state is
reported (max and RMS); so is a FAST step's (integer results exact, fields
within the same bounds):
  rho   against S_rho = m Poly6 sum_j [c^3 + 3 c^2 (h2 + r2)] (the sum's error
        scale: c = h2 - r2 cancels near h), and plainly relative to rho64
        over the bulk fluid (rho64 >= rho0/2) and over all particles
  p     against |p|_abs + |dp/drho| S_rho (the EOS terms cancel near rho0)
  dv    against the absolute-value evaluation S of the pair sum, and also
        plainly relative to |dv64| where |dv64| > 1e-3 S (cancellation-free)
  x     relative to |x64| after the whole step (input: the same FP32 state)
FP64 appears on the GPU only in the global reductions (mass, kinetic and
potential energy, max speed) that feed the CFL time step; those are checked
against a host float64 reduction of the read-back state, and a last step
with the CFL rule active (cfl 0.025: half the fixed step) is checked
against the oracle's step at the time step the rule picks.
The report is printed and, when the environment names a directory in
PIBITI_REPORT_DIR, also written there as precision_4m.json.
"""
import json
import os
from pathlib import Path

import numpy as np
import pytest

import oracle
from pibiti_amd import scenes, sph

pytestmark = pytest.mark.gpu


def field_errors(g, d, o32, p):
    rho0 = float(p.rest_density)
    out = {}

    def stats(e):
        e = np.asarray(e, np.float64)
        return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}

    dr = np.abs(g["dens"] - d["dens"])
    out["rho"] = stats(dr / d["dens_scale"].clip(1e-300))
    bulk = d["dens"] >= 0.5 * rho0
    out["rho_relative_bulk"] = stats(dr[bulk] / d["dens"][bulk])
    out["rho_relative_bulk"]["fraction_of_particles"] = float(bulk.mean())
    out["rho_relative_all"] = stats(dr / np.abs(d["dens"]).clip(1e-30))
    out["p"] = stats(np.abs(g["pres"] - d["pres"]) / d["pres_scale"])
    dd = np.abs(g["dvel"][:, :3] - d["dvel"][:, :3]).max(1)
    out["dv_scaled"] = stats(dd / d["dvel_scale"].clip(1e-300))
    mag = np.abs(d["dvel"][:, :3]).max(1)
    ok = mag > 1e-3 * d["dvel_scale"]
    out["dv_relative"] = stats(dd[ok] / mag[ok])
    out["dv_relative"]["fraction_of_particles"] = float(ok.mean())
    return out


def test_precision_study_4m_colliding_blocks():
    p, pos, vel, info = scenes.colliding_blocks(128, stiffness=400.0)
    n = int(p.num_particles)
    assert n == 4194304
    pre_steps = 2800  # the blocks start 0.24 apart closing at 2 m/s: contact after 2400 steps
    print("case: 4M colliding blocks, advance", pre_steps, "steps", flush=True)
    with sph.SphSystem(p) as s:
        s.reset(pos, vel)
        s.update(pre_steps)
        x0 = s.get_array(sph.SPH_ARRAY_POSITION)
        v0 = s.get_array(sph.SPH_ARRAY_VELOCITY)
        st = s.stats()
    # FP64 reductions on the device against the host
    m = float(p.particle_mass)
    vv = v0[:, :3].astype(np.float64)
    ke = 0.5 * m * np.sum(vv * vv)
    assert abs(st["kinetic_energy"] - ke) <= 1e-10 * ke
    assert abs(st["mass"] - m * n) <= 1e-12 * m * n
    assert abs(st["max_speed"] - np.sqrt((vv * vv).sum(1)).max()) <= 1e-12
    assert st["sort_errors"] == 0

    report = {"particles": n, "pre_steps": pre_steps, "stiffness": float(p.stiffness),
              "max_cell_count": st["max_cell_count"], "cap_cells": st["cap_cells"],
              "kinetic_energy": st["kinetic_energy"], "max_speed": st["max_speed"]}
    for prec, name in ((sph.SPH_PREC_REFERENCE, "reference"), (sph.SPH_PREC_KAHAN, "kahan")):
        q = p.copy()
        q.precision = prec
        print("case: 4M one step, precision", name, flush=True)
        with sph.SphSystem(q, keep_dvel=True) as s:
            s.reset(x0, v0)
            s.update(1)
            g = {k: s.get_array(i) for i, k in enumerate(["pos", "vel", "dens", "pres", "dvel"])}
        o = oracle.step_f32(q, x0, v0, prec=prec)
        for k in g:
            assert np.array_equal(g[k], o[k]), f"{name} {k} differs from the FP32 oracle"
        d = oracle.kernels_f64_on(o, q)
        errs = field_errors(g, d, o, q)
        w = oracle.step_f64(q, x0, v0)
        ex = np.abs(g["pos"][:, :3] - w["pos"][:, :3]).max(1) / np.abs(w["pos"][:, :3]).max(1)
        errs["x"] = {"max": float(ex.max()), "rms": float(np.sqrt(np.mean(ex * ex)))}
        report[name] = errs
        for f in ("rho", "rho_relative_bulk", "p", "dv_scaled", "x"):
            assert errs[f]["max"] < 1e-5, (name, f, errs[f])
    # compensation can only help the sums it compensates
    assert report["kahan"]["rho_relative_bulk"]["rms"] <= report["reference"]["rho_relative_bulk"]["rms"] * 1.05

    # FAST on the same state: integer results exact (keys, order, cell
    # ranges, force-pair counts; positions after the step), the fields under
    # the same 1e-5 bounds against FP64
    q = p.copy()
    q.precision = sph.SPH_PREC_FAST
    print("case: 4M one step, precision fast", flush=True)
    with sph.SphSystem(q, keep_dvel=True) as s:
        s.reset(x0, v0)
        s.update(1)
        g = {k: s.get_array(i) for i, k in enumerate(["pos", "vel", "dens", "pres", "dvel"])}
        nnb = s.get_array(sph.SPH_ARRAY_NEIGHBOURS)
        grid = s.get_grid()
    o = oracle.step_f32(p, x0, v0)
    for k in ("sorted_hash", "perm", "cell_start", "cell_end"):
        assert np.array_equal(grid[k], o[k]), f"fast {k}"
    assert np.array_equal(g["pos"], o["pos"]), "fast positions"
    cnt = oracle.neighbour_counts_f32(p, o["pos"][o["perm"]], o["cell_start"], o["cell_end"], pred=1)
    want = np.empty_like(cnt)
    want[o["perm"]] = cnt
    assert np.array_equal(nnb, want), "fast force pair counts"
    d = oracle.kernels_f64_on(o, p)
    errs = field_errors(g, d, o, p)
    report["fast"] = errs
    for f in ("rho", "rho_relative_bulk", "p", "dv_scaled"):
        assert errs[f]["max"] < 1e-5, ("fast", f, errs[f])

    # the CFL rule at this scale: the FP64 max speed of the 4M state picks the
    # step (half the fixed one here), and the step taken with it equals the
    # FP32 oracle's step at that time step bit for bit
    q = p.copy()
    q.cfl = 0.025
    vmax = float(np.sqrt((vv * vv).sum(1)).max())
    dt_expect = min(float(q.time_step), float(q.cfl) * float(q.h) / (float(q.sound_speed) + vmax))
    assert dt_expect < 0.75 * float(q.time_step)  # the rule is active
    print("case: 4M one CFL step", flush=True)
    with sph.SphSystem(q) as s:
        s.reset(x0, v0)
        assert abs(s.stats()["dt_cfl"] - dt_expect) <= 1e-15
        s.update(1)
        xc = s.get_array(sph.SPH_ARRAY_POSITION)
        vc = s.get_array(sph.SPH_ARRAY_VELOCITY)
        assert abs(s.stats()["sim_time"] - float(np.float32(dt_expect))) < 1e-15
    r = p.copy()
    r.time_step = np.float32(dt_expect)
    o = oracle.step_f32(r, x0, v0)
    assert np.array_equal(xc, o["pos"]) and np.array_equal(vc, o["vel"])
    report["cfl"] = {"cfl": float(q.cfl), "max_speed": vmax, "dt": dt_expect, "fixed_dt": float(q.time_step)}
    print("precision report:", json.dumps(report), flush=True)
    out_dir = os.environ.get("PIBITI_REPORT_DIR")
    if out_dir and Path(out_dir).is_dir():
        (Path(out_dir) / "precision_4m.json").write_text(json.dumps(report, indent=1))
