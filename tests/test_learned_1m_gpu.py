"""The learned surrogate at its workload: the 1,048,576-particle dam break
(128x64x128 lattice, spacing 0.01, h = 0.02, dt = 5e-5) stepped at 4 x dt.

A model is trained in the module fixture on 20 plain-solver trajectories
(the bench trains on 200), then:

* every surrogate step's densities and neighbour counts equal the oracle's
  for the same positions, bit for bit: the plain density kernel (reference
  order, cap included) judges each step;
* a tolerance nothing meets sends every step to the plain fallback, whose
  result is the plain solver's, bit for bit, at 1M;
* the ablation over the whole workload (1,000 dt of simulated time from the
  same state): the trained surrogate with the 1 % fallback, the plain step
  at 4 x dt (the method without the network: the pair force computed
  exactly, at the coarse step) and the zero network (no pair force at all),
  each against the plain run at dt: density error and simulated-time
  throughput.  The report is printed and, when the environment names a
  directory in PIBITI_REPORT_DIR, written there as learned_1m.json.
"""
import json
import os
import time
from pathlib import Path

import numpy as np
import pytest

import oracle
from pibiti_amd import learn, scenes, sph

pytestmark = pytest.mark.gpu

FACTOR = 4
WARM = 20


@pytest.fixture(scope="module")
def trained():
    p, pos, vel, _ = scenes.gpu_dam_break_1m()
    print("case: training 20 trajectories for the 1M surrogate", flush=True)
    model = sph.Model(learn.default_norm(p), seed=1)
    hist = learn.train(model, scenes_count=20, steps=800, stride=8, iters=32, lr=3e-3, lr_end=1e-4)
    assert hist["loss"][-1] < 0.5 * hist["loss"][0]
    with sph.SphSystem(p) as s:
        s.reset(pos, vel)
        s.update(WARM)
        x0 = s.get_array(sph.SPH_ARRAY_POSITION)
        v0 = s.get_array(sph.SPH_ARRAY_VELOCITY)
    yield p, x0, v0, model, hist
    model.close()


def system_at(p, x0, v0):
    """A context holding the state (x0, v0); callers take one plain step
    first (the surrogate starts from a stepped state)."""
    s = sph.SphSystem(p)
    s.reset(x0, v0)
    return s


def test_surrogate_steps_match_the_oracle_density_and_neighbours(trained):
    p, x0, v0, model, _ = trained
    n = int(p.num_particles)
    print("case: 1M surrogate steps against the oracle's density and neighbour counts", flush=True)
    with system_at(p, x0, v0) as s:
        s.update(1)
        for k in range(3):
            rep = s.step_learned(model, 1, factor=FACTOR, tol=1e9)
            assert rep["coarse_steps"] == 1 and rep["fallbacks"] == 0
            x = s.get_array(sph.SPH_ARRAY_POSITION)
            rho = s.get_array(sph.SPH_ARRAY_DENSITY)
            nn = s.get_array(sph.SPH_ARRAY_NEIGHBOURS)
            st = s.stats()
            assert st["sort_errors"] == 0
            g = oracle.grid_of(p, x)
            dens_sorted, _ = oracle.density_f32(p, g["spos"], g["cell_start"], g["cell_end"])
            cnt_sorted = oracle.neighbour_counts_f32(p, g["spos"], g["cell_start"], g["cell_end"], pred=0)
            dens = np.empty(n, np.float32)
            dens[g["perm"]] = dens_sorted
            cnt = np.empty(n, np.uint32)
            cnt[g["perm"]] = cnt_sorted
            bad = np.flatnonzero(rho != dens)
            assert bad.size == 0, f"step {k}: {bad.size} densities differ, first {bad[:5]}"
            bad = np.flatnonzero(nn != cnt)
            assert bad.size == 0, f"step {k}: {bad.size} neighbour counts differ, first {bad[:5]}"
            assert cnt.max() < 128  # no list overflow
            print(f"step {k}: cap cells {st['cap_cells']}, mean neighbours {cnt.mean():.2f}", flush=True)


def test_fallback_at_1m_replays_plain_steps_exactly(trained):
    p, x0, v0, model, _ = trained
    print("case: 1M surrogate with every step falling back", flush=True)
    with system_at(p, x0, v0) as a, system_at(p, x0, v0) as b:
        a.update(1)
        b.update(1)
        rep = a.step_learned(model, 2, factor=FACTOR, tol=1e-30)
        b.update(2 * FACTOR)
        assert rep["fallbacks"] == 2 and rep["plain_steps"] == 2 * FACTOR and rep["coarse_steps"] == 0
        for k in (sph.SPH_ARRAY_POSITION, sph.SPH_ARRAY_VELOCITY, sph.SPH_ARRAY_DENSITY):
            assert np.array_equal(a.get_array(k), b.get_array(k)), k
        assert abs(a.stats()["sim_time"] - b.stats()["sim_time"]) < 1e-12


def test_ablation_over_the_workload(trained):
    """1,000 dt of simulated time from the same state four ways; density
    error against the plain run at dt and simulated-time throughput."""
    p, x0, v0, model, hist = trained
    rho0 = float(p.rest_density)
    dt = float(p.time_step)
    coarse = 1000 // FACTOR

    def err(rho, ref):
        d = np.abs(np.asarray(rho, np.float64) - ref) / rho0
        return {"rms": float(np.sqrt(np.mean(d * d))), "max": float(d.max())}

    def timed(run):
        t0 = time.perf_counter()
        out = run()
        return out, time.perf_counter() - t0

    report = {"particles": int(p.num_particles), "dt": dt, "factor": FACTOR, "simulated_s": 1000 * dt,
              "training": {"trajectories": hist["scenes"], "steps": hist["steps"],
                           "first_loss": float(hist["loss"][0]), "final_loss": float(hist["loss"][-1])}}
    print("case: 1M ablation, plain run at dt", flush=True)
    with system_at(p, x0, v0) as s:
        s.update(1)
        s.synchronize()
        _, secs = timed(lambda: (s.update(1000), s.synchronize()))
        rho_ref = s.get_array(sph.SPH_ARRAY_DENSITY).astype(np.float64)
    report["plain_dt"] = {"seconds": secs, "sim_per_wall": 1000 * dt / secs}

    print("case: 1M ablation, trained surrogate with the 1% fallback", flush=True)
    with system_at(p, x0, v0) as s:
        s.update(1)
        s.step_learned(model, 0, factor=FACTOR, tol=0.01)  # buffers outside the timing
        s.synchronize()
        rep, secs = timed(lambda: s.step_learned(model, coarse, factor=FACTOR, tol=0.01))
        s.synchronize()
        sim = (FACTOR * rep["coarse_steps"] + rep["plain_steps"]) * dt
        assert abs(s.stats()["sim_time"] - (1 + 1000) * dt) < 1e-9
        report["trained"] = {"density_error": err(s.get_array(sph.SPH_ARRAY_DENSITY), rho_ref),
                             "seconds": secs, "sim_per_wall": sim / secs, "fallbacks": rep["fallbacks"],
                             "coarse_steps": rep["coarse_steps"]}
        assert s.stats()["sort_errors"] == 0

    print("case: 1M ablation, plain step at 4 dt", flush=True)
    q = p.copy()
    q.time_step = FACTOR * dt
    with system_at(p, x0, v0) as s:
        # the same first step at dt as the others, then the coarse steps
        s.update(1)
        s.set_params(q)
        s.synchronize()
        _, secs = timed(lambda: (s.update(coarse), s.synchronize()))
        report["plain_4dt"] = {"density_error": err(s.get_array(sph.SPH_ARRAY_DENSITY), rho_ref),
                               "seconds": secs, "sim_per_wall": 1000 * dt / secs}

    print("case: 1M ablation, zero network", flush=True)
    zero = sph.Model(learn.default_norm(p), seed=1)
    zero.set_weights(np.zeros(sph.MLP_NUM_PARAMS, np.float32))
    with system_at(p, x0, v0) as s:
        s.update(1)
        s.step_learned(zero, 0, factor=FACTOR, tol=1e9)
        s.synchronize()
        rep, secs = timed(lambda: s.step_learned(zero, coarse, factor=FACTOR, tol=1e9))
        s.synchronize()
        report["zero_network"] = {"density_error": err(s.get_array(sph.SPH_ARRAY_DENSITY), rho_ref),
                                  "seconds": secs, "sim_per_wall": 1000 * dt / secs}
    zero.close()

    for k in ("trained", "plain_4dt", "zero_network"):
        report[k]["speedup_vs_plain_dt"] = report[k]["sim_per_wall"] / report["plain_dt"]["sim_per_wall"]
    print("learned 1M report:", json.dumps(report), flush=True)
    out_dir = os.environ.get("PIBITI_REPORT_DIR")
    if out_dir:
        Path(out_dir).mkdir(parents=True, exist_ok=True)
        (Path(out_dir) / "learned_1m.json").write_text(json.dumps(report, indent=1))
    # what the network gains over no pair force at all
    assert report["trained"]["density_error"]["rms"] < 0.5 * report["zero_network"]["density_error"]["rms"]
    # every surrogate step is faster than four plain steps
    assert report["trained"]["speedup_vs_plain_dt"] > 1.0
