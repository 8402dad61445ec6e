"""Particle colours and dye (computeForceD's colouring, source/CUDA/System.cu:405-546).

CPU tests pin the oracle with closed forms over its own (separately pinned)
density, velocity and velocity increment: each colour mode's formula, the
hue switch, the foam potential of two particles, and the dye's pick-up,
fade and clear count.  GPU tests compare the device with the oracle bit for
bit over several steps.
"""
import math

import numpy as np
import pytest

import oracle
from pibiti_amd import sph

f32 = np.float32
BASE = np.array([0.2, 0.5, 1.0])


def params(n, clr, hue=0, dye_type=0, dye_clear=0, world=0.2, **kw):
    s = sph.scene_default()
    s.num_particles = n
    s.time_step = 1e-3
    for c in range(3):
        s.world_min[c] = -world
        s.world_max[c] = world
    s.particle_r = 0.005
    s.h = 0.02
    s.cell_size = 0.02
    s.particle_mass = 1e-3
    s.clr_type = clr
    s.hue = hue
    s.dye_type = dye_type
    s.dye_clear = dye_clear
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    return sph.params_from_scene(s)


def cluster(n=6, seed=1):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.012, 0.012, (n, 3))
    pos[:, 3] = 1
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.5, (n, 3))
    return pos, vel


def colour_step(p, pos, vel, dye=None):
    one = oracle.step_f32(p, pos, vel)  # density, integrated velocity, increment
    col = oracle.run_f32_colour(p, pos, vel, 1, dye=dye)
    return one, col


@pytest.mark.parametrize("hue", [0, 1])
@pytest.mark.parametrize("clr", [sph.SPH_CLR_DENS, sph.SPH_CLR_ACCEL, sph.SPH_CLR_DENS_ACC, sph.SPH_CLR_VEL_ACC,
                                 sph.SPH_CLR_VEL_RGB])
def test_colour_modes_closed_form(clr, hue):
    p = params(6, clr, hue=hue, brightness=0.35, contrast=0.4)
    pos, vel = cluster()
    one, out = colour_step(p, pos, vel)
    b, c, dt, rho0 = 0.35, float(f32(0.4)), float(f32(1e-3)), 1000.0
    for i in range(6):
        v = one["vel_int"][i, :3].astype(np.float64)
        a = one["dvel"][i, :3].astype(np.float64)
        rho = float(one["dens"][i])
        speed, acc = 2.5 * np.linalg.norm(v), 0.02 * np.linalg.norm(a) / dt
        dens = 4 * (rho - rho0) / rho0
        intens = 0.0
        if clr == sph.SPH_CLR_VEL_ACC:
            exp = BASE * (b + c * speed) + np.array([0.7, 0.35, 0]) * (c * acc)
        elif clr == sph.SPH_CLR_DENS_ACC:
            exp = BASE * (b + c * dens) + np.array([0.7, 0.7, 0]) * (c * acc)
        elif clr == sph.SPH_CLR_VEL_RGB:
            exp = 0.5 + c * 2.5 * v
        elif clr == sph.SPH_CLR_ACCEL:
            intens = b + c * acc
            exp = BASE * intens
        else:
            intens = b + c * dens
            exp = BASE * intens
        if hue == 1:
            exp = np.array([intens, 0.0, exp[2]])
        np.testing.assert_allclose(out["colour"][i, :3], exp, rtol=2e-5, atol=2e-6)
        assert out["colour"][i, 3] == 1


def test_no_colour_mode_writes_nothing():
    p = params(6, sph.SPH_CLR_NONE, dye_type=sph.SPH_DYE_BOX, dye_pos=[0, 0, 0], dye_size=[1, 1, 1])
    pos, vel = cluster()
    out = oracle.run_f32_colour(p, pos, vel, 2, dye=np.full(6, 0.5, np.float32))
    assert np.all(out["colour"] == 0) and np.all(out["dye"] == f32(0.5))


def test_foam_potential_two_particles():
    """Nd = clamp01((KE - 0.1)/0.9) 3 clamp01((vdiff - 5)/15), vdiff =
    |v1 - v2| (1 - n_v . n_x) (1 - r)/3 with n_v, n_x the normalised float4
    sums (w included) of the velocities and positions."""
    p = params(2, sph.SPH_CLR_VEL, particle_mass=0.2)
    pos = np.array([[0.001, 0.002, 0.0, 1.0], [0.004, -0.003, 0.002, 1.0]], np.float32)
    vel = np.array([[9.0, -4.0, 2.0, 0.0], [-8.0, 6.0, 1.0, 0.0]], np.float32)
    one, out = colour_step(p, pos, vel)
    x = one["pos"].astype(np.float64)
    v = one["vel_int"].astype(np.float64)
    m = float(f32(0.2))
    for i, j in ((0, 1), (1, 0)):
        nv = (v[i] + v[j]) / np.linalg.norm(v[i] + v[j])
        nx = (x[i] + x[j]) / np.linalg.norm(x[i] + x[j])
        r = np.linalg.norm(x[i] - x[j])
        vdiff = np.linalg.norm(v[i] - v[j]) * (1 - nv @ nx) * (1 - r) / 3
        clamp = lambda t, lo, hi: (min(t, hi) - min(t, lo)) / (hi - lo)
        ke = m * 1000 * (v[i] @ v[i]) * 0.5
        nd = clamp(ke, 0.1, 1.0) * 3 * clamp(vdiff, 5.0, 20.0)
        assert vdiff > 5 and 0 < nd
        np.testing.assert_allclose(out["colour"][i, :3], [nd] * 3, rtol=1e-5)


def test_dye_pick_up_fade_and_clear():
    """Inside the dye box a particle's dye becomes 1 - fade dt; outside it
    fades by fade dt a step, floored at 0; while the clear count (counted
    down before each step) is above 0 the dye is cleared instead."""
    n = 4
    pos = np.zeros((n, 4), np.float32)
    pos[:, 3] = 1
    pos[:, 0] = [0.0, 0.05, 0.1, 0.15]
    vel = np.zeros((n, 4), np.float32)
    kw = dict(dye_type=sph.SPH_DYE_BOX, dye_pos=[0.0, 0.0, 0.0], dye_size=[0.02, 0.02, 0.02], dye_fade=50.0,
              gravity=[0, 0, 0])
    p = params(n, sph.SPH_CLR_DENS, **kw)
    start = np.array([0.3, 0.3, 0.01, 0.0], np.float32)
    out = oracle.run_f32_colour(p, pos, vel, 1, dye=start)
    step = float(f32(f32(1e-3) * f32(50.0)))
    np.testing.assert_allclose(out["dye"], [1 - step, 0.3 - step, 0.0, 0.0], atol=1e-7)
    base = out["colour"][:, :3] - np.outer(out["dye"], [0.9, 0.9, 1.0])
    assert np.allclose(base / BASE, (base / BASE)[:, :1], atol=1e-5)  # dye added on top of the mode colour
    # a clear count of 2: the first step sees 1 and clears, the second applies the dye
    p = params(n, sph.SPH_CLR_DENS, dye_clear=2, **kw)
    out = oracle.run_f32_colour(p, pos, vel, 1, dye=start)
    np.testing.assert_array_equal(out["dye"], 0)
    out = oracle.run_f32_colour(p, pos, vel, 2, dye=start)
    np.testing.assert_allclose(out["dye"][0], 1 - step, atol=1e-7)
    # sphere of radius dye_size[1]
    p = params(n, sph.SPH_CLR_DENS, **{**kw, "dye_type": sph.SPH_DYE_SPHERE, "dye_size": [0.0, 0.06, 0.0]})
    out = oracle.run_f32_colour(p, pos, vel, 1, dye=np.zeros(n, np.float32))
    np.testing.assert_allclose(out["dye"], [1 - step, 1 - step, 0, 0], atol=1e-7)


# ---------------------------------------------------------------- GPU parity
GPU_CASES = [(sph.SPH_CLR_VEL_ACC, 0, sph.SPH_DYE_BOX, 2), (sph.SPH_CLR_DENS_ACC, 1, sph.SPH_DYE_SPHERE, 0),
             (sph.SPH_CLR_VEL, 0, 0, 0), (sph.SPH_CLR_VEL, 1, sph.SPH_DYE_BOX, 0),
             (sph.SPH_CLR_VEL_RGB, 0, sph.SPH_DYE_SPHERE, 1), (sph.SPH_CLR_ACCEL, 1, 0, 0),
             (sph.SPH_CLR_DENS, 0, sph.SPH_DYE_BOX, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES)
def test_colours_bit_exact_on_gpu(case):
    clr, hue, dye_type, clear = case
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(*[np.arange(-0.1, 0.1, 0.01)] * 3, indexing="ij"), -1).reshape(-1, 3)
    n = len(g)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = g + rng.uniform(-0.001, 0.001, g.shape)
    pos[:, 3] = 1
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1.0, (n, 3))
    p = params(n, clr, hue=hue, dye_type=dye_type, dye_clear=clear, dye_pos=[0.02, -0.03, 0.0],
               dye_size=[0.03, 0.04, 0.05], dye_fade=20.0, time_step=2e-4, particle_mass=1e-3,
               collider=1, coll_pos=[0.05, 0.0, 0.0], coll_r=0.02)
    steps = 6
    print("case: colour", case, "n =", n, flush=True)
    with sph.SphSystem(p) as s:
        s.reset(pos, vel)
        s.update(steps)
        got = dict(pos=s.get_array(sph.SPH_ARRAY_POSITION), vel=s.get_array(sph.SPH_ARRAY_VELOCITY),
                   colour=s.get_array(sph.SPH_ARRAY_COLOR), dye=s.get_array(sph.SPH_ARRAY_DYE))
    o = oracle.run_f32_colour(p, pos, vel, steps)
    for k in ("pos", "vel", "colour", "dye"):
        assert np.array_equal(got[k], o[k]), f"{case} {k}: max diff {np.abs(got[k] - o[k]).max()}"
    if dye_type:
        assert o["dye"].max() > 0


FAST_CASES = [(sph.SPH_CLR_VEL, 0, sph.SPH_DYE_BOX, 0), (sph.SPH_CLR_DENS_ACC, 1, sph.SPH_DYE_SPHERE, 0),
              (sph.SPH_CLR_ACCEL, 1, 0, 0)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", FAST_CASES)
def test_colours_fast_mode_one_step(case):
    """The colour pass after a FAST step: the colours follow the step's
    velocities, densities and force increments, which FAST keeps within
    1e-5 of the reference, so the colours agree within 1e-4 (absolute,
    on components in [0, 1]); positions and the dye, which follows them,
    are exact."""
    clr, hue, dye_type, clear = case
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(*[np.arange(-0.1, 0.1, 0.01)] * 3, indexing="ij"), -1).reshape(-1, 3)
    n = len(g)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = g + rng.uniform(-0.001, 0.001, g.shape)
    pos[:, 3] = 1
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 1.0, (n, 3))
    p = params(n, clr, hue=hue, dye_type=dye_type, dye_clear=clear, dye_pos=[0.02, -0.03, 0.0],
               dye_size=[0.03, 0.04, 0.05], dye_fade=20.0, time_step=2e-4, particle_mass=1e-3,
               collider=1, coll_pos=[0.05, 0.0, 0.0], coll_r=0.02)
    q = p.copy()
    q.precision = sph.SPH_PREC_FAST
    print("case: FAST colour", case, "n =", n, flush=True)
    with sph.SphSystem(q) as s:
        s.reset(pos, vel)
        s.update(1)
        got = dict(pos=s.get_array(sph.SPH_ARRAY_POSITION), colour=s.get_array(sph.SPH_ARRAY_COLOR),
                   dye=s.get_array(sph.SPH_ARRAY_DYE))
    o = oracle.run_f32_colour(p, pos, vel, 1)
    assert np.array_equal(got["pos"], o["pos"])
    assert np.array_equal(got["dye"], o["dye"])
    err = float(np.abs(got["colour"].astype(np.float64) - o["colour"]).max())
    print(case, "colour max abs diff", err, flush=True)
    assert err < 1e-4, f"{case}: colour differs by {err}"


def hip_runtime():
    import ctypes as C
    h = C.CDLL("libamdhip64.so")
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h


@pytest.mark.gpu
def test_device_view_reads_the_state_in_place():
    """The device view's slots, reordered by its index, are the arrays the
    copies return (the renderer's zero-copy path)."""
    import ctypes as C
    pos, vel = cluster(500, seed=4)
    pos[:, :3] *= 8
    p = params(500, sph.SPH_CLR_VEL_ACC, time_step=2e-4)
    print("case: device view", flush=True)
    with sph.SphSystem(p) as s:
        s.reset(pos, vel)
        s.update(3)
        v = s.device_view()
        assert v.n == 500 and v.pos4 and v.vel4 and v.colour4 and v.index
        hip = hip_runtime()
        assert hip.hipStreamSynchronize(v.stream) == 0
        got = {}
        for name, ptr, rec, dt in (("pos", v.pos4, 16, np.float32), ("vel", v.vel4, 16, np.float32),
                                   ("colour", v.colour4, 16, np.float32), ("index", v.index, 4, np.uint32)):
            buf = np.empty(500 * rec // np.dtype(dt).itemsize, dt)
            assert hip.hipMemcpy(buf.ctypes.data_as(C.c_void_p), ptr, 500 * rec, 2) == 0
            got[name] = buf
        idx = got["index"]
        assert sorted(idx.tolist()) == list(range(500))
        for name, which in (("pos", sph.SPH_ARRAY_POSITION), ("vel", sph.SPH_ARRAY_VELOCITY),
                            ("colour", sph.SPH_ARRAY_COLOR)):
            orig = np.empty((500, 4), np.float32)
            orig[idx] = got[name].reshape(500, 4)
            assert np.array_equal(orig, s.get_array(which)), name
