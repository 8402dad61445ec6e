"""The FAST precision mode against the oracle, under the task's float
contract: integer outputs exact (cell keys, sort permutation, cell ranges,
and each particle's number of force pairs, i.e. the neighbour sets of the
capped 27-cell walk), and every float within 1e-5 of both the FP64 oracle
and the FP32 restatement of the reference, each error scaled by the
condition of the value (the scales of test_gpu_parity.py).  Positions after
one step are bit-exact: the integration is the reference-order kernel; the
densities, pressures and force increments of that step are not.

FAST is deterministic (a fixed summation order per particle): a batch
equals its scenes run alone bit for bit, and graph replays equal eager
steps bit for bit.
"""
import numpy as np
import pytest

import oracle
from conftest import scaled_err
from pibiti_amd import scenes, sph

pytestmark = pytest.mark.gpu

TOL = 1e-5
FAST = sph.SPH_PREC_FAST


def read_state(s):
    return dict(pos=s.get_array(sph.SPH_ARRAY_POSITION), vel=s.get_array(sph.SPH_ARRAY_VELOCITY),
                dens=s.get_array(sph.SPH_ARRAY_DENSITY), pres=s.get_array(sph.SPH_ARRAY_PRESSURE),
                dvel=s.get_array(sph.SPH_ARRAY_DVEL), nnb=s.get_array(sph.SPH_ARRAY_NEIGHBOURS),
                grid=s.get_grid(), stats=s.stats())


def fast_step(p, pos, vel):
    q = p.copy()
    q.precision = FAST
    with sph.SphSystem(q, keep_dvel=True) as s:
        s.reset(pos, vel)
        s.update(1)
        return read_state(s)


def check_fast_step(p, pos, vel, what):
    return compare_fast(fast_step(p, pos, vel), p, pos, vel, what)


def compare_fast(g, p, pos, vel, what):
    """g: the state after one FAST step from (pos, vel), checked against the
    oracle's step from the same state under the FAST contract."""
    o = oracle.step_f32(p, pos, vel)
    gr = g["grid"]
    for k in ("sorted_hash", "perm", "cell_start", "cell_end"):
        assert np.array_equal(gr[k], o[k]), f"{what} {k}"
    assert np.array_equal(g["pos"], o["pos"]), f"{what} positions"
    # force pairs per particle: the capped walk's candidates with r2 <= force_r2_max
    cnt = oracle.neighbour_counts_f32(p, o["pos"][o["perm"]], o["cell_start"], o["cell_end"], pred=1)
    nnb = np.empty_like(cnt)
    nnb[o["perm"]] = cnt
    assert np.array_equal(g["nnb"], nnb), f"{what} force pair counts"
    d = oracle.kernels_f64_on(o, p)
    errs = {}
    for ref_name, ref in (("f64", d), ("f32", o)):
        errs[ref_name] = dict(
            dens=scaled_err(g["dens"], ref["dens"], d["dens_scale"]),
            pres=scaled_err(g["pres"], ref["pres"], d["pres_scale"]),
            dvel=scaled_err(g["dvel"][:, :3], ref["dvel"][:, :3], d["dvel_scale"]))
    bulk = d["dens"] >= 0.5 * float(p.rest_density)
    errs["f64"]["dens_bulk_rel"] = scaled_err(g["dens"][bulk], d["dens"][bulk], d["dens"][bulk])
    w = oracle.step_f64(p, pos, vel)
    errs["f64"]["pos"] = scaled_err(g["pos"][:, :3], w["pos"][:, :3], np.abs(w["pos"][:, :3]).max(1))
    # velocities after the update: relative to the largest speed of the step
    vs = np.abs(o["vel"][:, :3]).max()
    errs["f32"]["vel"] = float(np.abs(g["vel"][:, :3].astype(np.float64) - o["vel"][:, :3]).max() / vs)
    print(what, "scaled errors:", errs, flush=True)
    for ref_name in errs:
        for k, e in errs[ref_name].items():
            assert e < TOL, f"{what} {k} vs {ref_name}: {e}"
    return g


def test_fast_single_step_15k():
    p, pos, vel, _ = scenes.cpu_dam_break_15k()
    print("case: FAST 15k single step", flush=True)
    check_fast_step(p, pos, vel, "15k")


def test_fast_single_step_1m():
    p, pos, vel, _ = scenes.gpu_dam_break_1m()
    print("case: FAST 1M single step", flush=True)
    check_fast_step(p, pos, vel, "1M")


def test_fast_single_step_1m_with_crowded_cells():
    """After 200 reference steps, 40 particles are moved into two cells (20
    each, spread over the cell), so that those cells hold more than
    max_par_in_cell particles and the cap truncates their candidates."""
    p, pos, vel, _ = scenes.gpu_dam_break_1m()
    print("case: FAST 1M single step with crowded cells", flush=True)
    with sph.SphSystem(p) as s:
        s.reset(pos, vel)
        s.update(200)
        x = s.get_array(sph.SPH_ARRAY_POSITION)
        v = s.get_array(sph.SPH_ARRAY_VELOCITY)
    rng = np.random.default_rng(7)
    cs = float(p.cell_size[0])
    for k, corner in enumerate(((0.50, 0.30, 0.50), (0.81, 0.21, 0.33))):
        lo = np.floor(np.asarray(corner) / cs) * cs
        idx = np.arange(1000 + 20 * k, 1020 + 20 * k) * 37
        x[idx, :3] = (lo + cs * rng.uniform(0.1, 0.9, (20, 3))).astype(np.float32)
    o = oracle.step_f32(p, x, v)
    assert o["counters"][0] >= 2  # cells above the cap
    check_fast_step(p, x, v, "1M crowded")


def test_fast_partial_upload_between_steps():
    """Host uploads between FAST steps (the emitters' path): velocities of 300
    particles and positions of 100 replaced after three steps.  The next step
    starts from the uploaded state, so its order, packed candidate positions
    and walk see the new positions: the FAST contract against the oracle's
    step from that state (the reference-mode twin is
    test_gpu_parity.py::test_partial_upload_between_steps)."""
    print("case: FAST partial upload between steps", flush=True)
    p, pos, vel, _ = scenes.scaled_dam_break((20, 12, 20), seed=4)
    q = p.copy()
    q.precision = FAST
    with sph.SphSystem(q, keep_dvel=True) as s:
        s.reset(pos, vel)
        s.update(3)
        x = s.get_array(sph.SPH_ARRAY_POSITION)
        v = s.get_array(sph.SPH_ARRAY_VELOCITY)
        newv = v.copy()
        newv[1000:1300, :3] = np.float32([0.0, -0.5, 0.1])
        s.set_array(sph.SPH_ARRAY_VELOCITY, newv[1000:1300], first=1000)
        newx = x.copy()
        newx[2000:2100, 1] += np.float32(0.05)
        s.set_array(sph.SPH_ARRAY_POSITION, newx[2000:2100], first=2000)
        s.update(1)
        g = read_state(s)
    compare_fast(g, p, newx, newv, "FAST partial upload")


def test_fast_after_reference_steps():
    """A context switched from the reference order to FAST between steps
    (sph_set_params): the FAST step after the switch meets the FAST contract
    from the state the reference steps left, and switching back gives the
    reference step bit for bit."""
    print("case: reference steps, then FAST, then reference again", flush=True)
    p, pos, vel, _ = scenes.scaled_dam_break((20, 12, 20), seed=8)
    with sph.SphSystem(p, keep_dvel=True) as s:
        s.reset(pos, vel)
        s.update(2)
        x = s.get_array(sph.SPH_ARRAY_POSITION)
        v = s.get_array(sph.SPH_ARRAY_VELOCITY)
        q = p.copy()
        q.precision = FAST
        s.set_params(q)
        s.update(1)
        g = read_state(s)
        s.set_params(p)
        s.update(1)
        r = read_state(s)
    compare_fast(g, p, x, v, "FAST after reference steps")
    o = oracle.step_f32(p, g["pos"], g["vel"])
    for k in ("pos", "vel", "dens", "pres"):
        assert np.array_equal(r[k], o[k]), f"reference step after FAST: {k}"
    cnt = oracle.neighbour_counts_f32(p, o["pos"][o["perm"]], o["cell_start"], o["cell_end"], pred=1)
    want = np.empty_like(cnt)
    want[o["perm"]] = cnt
    assert np.array_equal(r["nnb"], want), "reference step after FAST: force pair counts"


def test_fast_batch_equals_scenes_alone_and_graph_replays():
    lat = [(10, 8, 10), (14, 6, 9), (7, 12, 7)]
    sc = [scenes.scaled_dam_break(l, seed=40 + k, vel_sigma=0.5, precision=FAST) for k, l in enumerate(lat)]
    print("case: FAST batch of 3 scenes, 6 eager steps + 3 replays of 2", flush=True)
    alone = []
    for q, x, v, _ in sc:
        with sph.SphSystem(q) as s:
            s.reset(x, v)
            s.update(12)
            alone.append([s.get_array(k) for k in (sph.SPH_ARRAY_POSITION, sph.SPH_ARRAY_VELOCITY,
                                                    sph.SPH_ARRAY_DENSITY)])
    with sph.SphBatch([q for q, _, _, _ in sc]) as b:
        for k, (_, x, v, _) in enumerate(sc):
            b.reset(k, x, v)
        b.update(6)
        b.capture(2)
        b.replay(3)
        for k in range(len(sc)):
            got = [b.get_array(k, w) for w in (sph.SPH_ARRAY_POSITION, sph.SPH_ARRAY_VELOCITY,
                                                sph.SPH_ARRAY_DENSITY)]
            for a, c in zip(got, alone[k]):
                assert np.array_equal(a, c), f"scene {k}"


def test_fast_rejects_large_caps():
    p, _, _, _ = scenes.scaled_dam_break((6, 6, 6), seed=1, precision=FAST)
    q = p.copy()
    q.max_par_in_cell = 65
    with pytest.raises(sph.SphError) as ex:
        sph.SphSystem(q)
    assert ex.value.status == 7  # SPH_ERR_UNSUPPORTED


def test_fast_long_run_tracks_the_reference_mode():
    """1,000 steps of the 15,625-particle dam break in both modes: the float
    contract holds per step, and over the run the two trajectories stay
    together (measured: positions within 7.6e-4 h at the worst particle and
    6e-6 h RMS, kinetic energy within 1.3e-7, mean density within 2.4e-7,
    the same top speed); the bounds asserted are ten or more times wider."""
    res = {}
    for prec in (sph.SPH_PREC_REFERENCE, FAST):
        p, pos, vel, _ = scenes.cpu_dam_break_15k(precision=prec)
        print("case: 1000 steps, precision", prec, flush=True)
        with sph.SphSystem(p) as s:
            s.reset(pos, vel)
            s.update(1000)
            res[prec] = (s.get_array(sph.SPH_ARRAY_POSITION), s.get_array(sph.SPH_ARRAY_DENSITY), s.stats())
    (xa, da, sa), (xb, db, sb) = res[sph.SPH_PREC_REFERENCE], res[FAST]
    h = float(p.h)
    dx = np.abs(xa[:, :3].astype(np.float64) - xb[:, :3]).max(axis=1) / h
    print("max |dx|/h", dx.max(), "rms", np.sqrt(np.mean(dx * dx)), flush=True)
    assert np.isfinite(xb).all() and np.isfinite(db).all()
    assert dx.max() < 1e-2 and np.sqrt(np.mean(dx * dx)) < 1e-4
    assert abs(sa["kinetic_energy"] - sb["kinetic_energy"]) < 1e-5 * sa["kinetic_energy"]
    assert abs(float(da.mean()) - float(db.mean())) < 1e-5 * float(da.mean())
    assert abs(sa["max_speed"] - sb["max_speed"]) < 1e-4 * sa["max_speed"]
    assert sa["sort_errors"] == sb["sort_errors"] == 0
