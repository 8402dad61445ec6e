"""Pin the oracle's boundary shapes, z effects, collider and accelerators
(source/CUDA/System.cu:41-159, :374-399; source/CUDA/Kernel_Cell.cui:78-96)
with closed forms evaluated independently in float64 numpy.

Each case puts one particle where a known set of walls acts and checks the
velocity and position after the oracle's integrate (or, for the force-side
terms, the velocity increment of a one-particle step) against the walls
applied by hand in the reference's order.  Mistakes these catch: a wall of
the wrong shape acting (a box wall inside a cylinder scene), the wrong
damping constant (bndDamp for addB walls, bndDampC for addC walls), walls in
the wrong order, a missing or doubled wrap, the wave wall not moving with
the phase, the phase advanced after instead of before the step, and the
collider's normal, spring, damping and shear terms.
"""
import math

import numpy as np
import pytest

import oracle
from pibiti_amd import sph

R = 0.005          # particle radius: soft band 8 R = 0.04, hard clamp 1 R
DT = 1e-4
EPS = 1e-5         # EPS of boundary()


def shape_params(bnd=sph.SPH_BND_BOX, eff=sph.SPH_BND_EFF_NONE, n=1, world=0.2, **kw):
    s = sph.scene_default()
    s.num_particles = n
    s.time_step = DT
    w = world if isinstance(world, tuple) else (world,) * 3
    for c in range(3):
        s.gravity[c] = 0.0
        s.world_min[c] = -w[c]
        s.world_max[c] = w[c]
    s.particle_r = R
    s.h = 0.02
    s.cell_size = 0.02
    s.particle_mass = 1e-3
    s.bnd_type = bnd
    s.bnd_eff_z = eff
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    return sph.params_from_scene(s)


def one(x, v):
    pos = np.zeros((1, 4), np.float32)
    vel = np.zeros((1, 4), np.float32)
    pos[0, :3] = x
    pos[0, 3] = 1
    vel[0, :3] = v
    return pos, vel


def f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def wall(v, diff, n, stiff, damp):
    """addB/addC: acc = stiff*diff - damp*(n.v); v += acc*n*dt (when diff > EPS)."""
    if diff <= EPS:
        return v
    n = np.asarray(n, np.float64)
    return v + (stiff * diff - damp * float(n @ v)) * n * DT


def integrate(p, x, v):
    pos, vel = one(x, v)
    x1, v1 = oracle.integrate_f32(p, pos, vel)
    return x1[0, :3].astype(np.float64), v1[0, :3].astype(np.float64)


def check(p, x0, v0, v_exp, x_shift=(0.0, 0.0, 0.0)):
    x1, v1 = integrate(p, x0, v0)
    np.testing.assert_allclose(v1, v_exp, rtol=0, atol=2e-6)
    x_exp = f32(x0) + np.asarray(x_shift) + v_exp * DT
    np.testing.assert_allclose(x1, x_exp, rtol=0, atol=1e-7)


def consts(p):
    return float(p.dist_bnd_soft), float(p.bnd_stiff), float(p.bnd_damp), float(p.bnd_damp_c)


# ------------------------------------------------------------------ shapes
def test_sphere_wall_only():
    """Inside the sphere's band near +x: the sphere wall acts (bndDampC,
    normal -x/|x|); the box walls, which would also reach here, do not."""
    p = shape_params(sph.SPH_BND_SPHERE)
    b, st, dp, dc = consts(p)
    x0 = f32([0.17, 0.02, -0.03])
    v0 = f32([0.3, -0.2, 0.1])
    ln = np.linalg.norm(x0)
    v = wall(v0, b + ln - 0.2, -x0 / ln, st, dc)
    check(p, x0, v0, v)


def test_cylinder_y_keeps_y_walls_drops_x_z_walls():
    """Cylinder about y (radius world_max.x): the y- wall (bndDamp) acts
    first, then the cylinder (bndDampC); the x+ wall that a box would apply
    here does not."""
    p = shape_params(sph.SPH_BND_CYL_Y)
    b, st, dp, dc = consts(p)
    x0 = f32([0.17, -0.17, 0.0])
    v0 = f32([0.25, -0.3, 0.05])
    v = wall(v0, (b - x0[1]) - 0.2, [0, 1, 0], st, dp)
    ln = math.hypot(x0[0], x0[2])
    v = wall(v, b + ln - 0.2, [-x0[0] / ln, 0, -x0[2] / ln], st, dc)
    check(p, x0, v0, v)


def test_cylinder_z_keeps_z_walls_drops_x_y_walls():
    """Cylinder about z (radius -world_min.y): the z+ wall, then the cylinder."""
    p = shape_params(sph.SPH_BND_CYL_Z)
    b, st, dp, dc = consts(p)
    x0 = f32([0.17, 0.0, 0.17])
    v0 = f32([0.2, 0.1, 0.3])
    v = wall(v0, (b + x0[2]) - 0.2, [0, 0, -1], st, dc)
    ln = math.hypot(x0[0], x0[1])
    v = wall(v, b + ln - 0.2, [-x0[0] / ln, -x0[1] / ln, 0], st, dc)
    check(p, x0, v0, v)


def test_cylinders_yz_keep_the_box():
    """Both cylinders plus every box wall: x+ (bndDamp), then cylinder Y,
    then cylinder Z (both bndDampC)."""
    p = shape_params(sph.SPH_BND_CYL_YZ)
    b, st, dp, dc = consts(p)
    x0 = f32([0.17, 0.0, 0.0])
    v0 = f32([0.3, 0.05, -0.02])
    v = wall(v0, (b + x0[0]) - 0.2, [-1, 0, 0], st, dp)
    ln = math.hypot(x0[0], x0[2])
    v = wall(v, b + ln - 0.2, [-x0[0] / ln, 0, -x0[2] / ln], st, dc)
    ln = math.hypot(x0[0], x0[1])
    v = wall(v, b + ln - 0.2, [-x0[0] / ln, -x0[1] / ln, 0], st, dc)
    check(p, x0, v0, v)


# --------------------------------------------------------------- z effects
def span(p):
    b = float(p.dist_bnd_soft)
    return ((0.4 - 2 * b) - R)


def test_cycle_wraps_both_ways_without_z_walls():
    p = shape_params(eff=sph.SPH_BND_EFF_CYCLE)
    b = float(p.dist_bnd_soft)
    ev = float(p.exit_vel)
    # leaving through z+ faster than exit_vel: moved back by the span
    x0, v0 = f32([0.0, 0.0, 0.2 - b + 0.001]), f32([0.0, 0.0, 2 * ev])
    check(p, x0, v0, v0, x_shift=(0, 0, -span(p)))
    # leaving through z-: moved forward
    x0, v0 = f32([0.0, 0.0, -0.2 + b - 0.001]), f32([0.0, 0.0, -2 * ev])
    check(p, x0, v0, v0, x_shift=(0, 0, span(p)))
    # slower than exit_vel: no wrap, and no z wall either
    x0, v0 = f32([0.0, 0.0, 0.2 - b + 0.001]), f32([0.0, 0.0, 0.5 * ev])
    check(p, x0, v0, v0)


def test_wrap_is_one_way_and_keeps_the_far_wall():
    p = shape_params(eff=sph.SPH_BND_EFF_WRAP)
    b, st, dp, dc = consts(p)
    ev = float(p.exit_vel)
    x0, v0 = f32([0.0, 0.0, -0.2 + b - 0.001]), f32([0.0, 0.0, -2 * ev])
    check(p, x0, v0, v0, x_shift=(0, 0, span(p)))
    # at z+ the wall acts and nothing wraps
    x0, v0 = f32([0.0, 0.0, 0.2 - b + 0.001]), f32([0.0, 0.0, 2 * ev])
    v = wall(v0, (b + x0[2]) - 0.2, [0, 0, -1], st, dc)
    check(p, x0, v0, v)


@pytest.mark.parametrize("phase", [0.3, -1.2, 2.5])
def test_wave_floor_and_moving_z_wall(phase):
    """Wave effect: the sloped floor (bndDamp, normal (0, 1-sl, sl) with
    sl = -slope) and the z- wall moved by ampl (1 + sin(phase)) (bndDampC)."""
    slope, ampl = 0.2, 0.05
    p = shape_params(eff=sph.SPH_BND_EFF_WAVE, rot2_angle=slope, rot_twist=ampl, rot_angle=phase)
    b, st, dp, dc = consts(p)
    shift = float(np.float32(ampl)) * (1 + math.sin(float(np.float32(phase))))
    # float rounding of sin (|sin| <= 1) and of 1 + sin: a few ulp of ampl
    assert float(p.wave_zmin_shift) == pytest.approx(shift, abs=4 * 2.0**-24 * ampl)
    sl = -float(np.float32(slope))
    z = -0.2 + shift + 0.5 * b
    x0 = f32([0.0, -0.2 + 0.02, z])
    v0 = f32([0.0, -0.1, -0.2])
    x0f = f32(x0)
    v = wall(v0, (b - (x0f[1] + 0.2)) - (x0f[2] + 0.2) * sl, [0, 1 - sl, sl], st, dp)
    v = wall(v, (b - x0f[2]) + (-0.2 + float(p.wave_zmin_shift)), [0, 0, 1], st, dc)
    v = wall(v, (b - x0f[1]) - 0.2, [0, 1, 0], st, dp)
    check(p, x0, v0, v)


def test_wave_phase_advances_before_each_step():
    """A step of a moving wave equals a step of a still wave whose phase is
    one advance ahead (App::UpdateEmitter runs before cSPH::Update)."""
    speed = 6.0
    kw = dict(eff=sph.SPH_BND_EFF_WAVE, rot_twist=0.05, rot2_angle=0.1, n=2)
    moving = shape_params(rot_angle=-1.5707963, rot_vel=speed, **kw)
    ahead = shape_params(rot_angle=float(np.float32(np.float32(-1.5707963) + np.float32(speed) * np.float32(DT))),
                         **kw)
    b = float(moving.dist_bnd_soft)
    pos = np.zeros((2, 4), np.float32)
    pos[:, 3] = 1
    pos[0, :3] = (0.0, 0.0, -0.2 + b * 0.3)
    pos[1, :3] = (0.05, 0.0, -0.2 + b * 0.6)
    vel = np.zeros((2, 4), np.float32)
    vel[:, 2] = -0.2
    a = oracle.run_f32(moving, pos, vel, 1)
    c = oracle.run_f32(ahead, pos, vel, 1)
    np.testing.assert_array_equal(a["vel"], c["vel"])
    np.testing.assert_array_equal(a["pos"], c["pos"])
    # and it did move: a still wave at the old phase differs
    still = shape_params(rot_angle=-1.5707963, **kw)
    d = oracle.run_f32(still, pos, vel, 1)
    assert not np.array_equal(a["vel"], d["vel"])


# -------------------------------------------------------------------- pump
PUMP_WORLD = (0.18, 0.3, 0.15)   # half extents of the reference's pump scenes


def pump(**kw):
    return shape_params(sph.SPH_BND_PUMP_Y, world=PUMP_WORLD, pump_ang_out=0.15,
                        shape_s=[0.9, 0.9, 2.01, 0.8585, 0.9, 0.9], **kw)


def test_pump_frame_outside_the_outlet():
    p = pump()
    b, st, dp, dc = consts(p)
    rad = 0.18
    x0 = f32([0.09, -0.13, -0.1])      # y < 0: frame wall
    v0 = f32([0.2, -0.3, 0.0])
    ln = math.hypot(x0[0], x0[1])
    assert b + ln - rad > EPS
    v = wall(v0, (b + ln) - rad, [-x0[0] / ln, -x0[1] / ln, 0], st, dp)
    check(p, x0, v0, v)


def test_pump_frame_open_over_the_outlet():
    p = pump()
    b, st, dp, dc = consts(p)
    x0 = f32([0.005, 0.16, -0.1])      # |atan(x/y)| < ang_out, y > 0: no frame wall
    v0 = f32([0.0, 0.1, 0.0])
    assert abs(math.atan(x0[0] / x0[1])) < 0.15
    check(p, x0, v0, v0)


def test_pump_lid():
    p = pump()
    b, st, dp, dc = consts(p)
    hc = float(p.h_close)
    x0 = f32([0.0, 0.0, hc - 0.95 * b])  # below hc - b s2 (s2 = 0.9), within b of hc
    v0 = f32([0.0, 0.0, 0.1])
    v = wall(v0, (b + x0[2]) - hc, [0, 0, -1], st, dp)
    check(p, x0, v0, v)


def test_pump_outlet_maps_onto_the_inlet():
    """Leaving through the outlet (y within exit_dist of the top, vy > exit_vel)
    re-enters on the inlet circle at z = wmax.z - b - |vy - exit_vel| 0.01 with
    (vx, vy, vz) -> (vx, vz, -vy).  The y+ box wall acts first."""
    p = pump()
    b, st, dp, dc = consts(p)
    hc = float(p.h_close)
    rad, wy, wz = PUMP_WORLD
    ev = float(p.exit_vel)
    x0 = f32([0.01, wy - 0.01, -0.1])
    v0 = f32([0.05, 0.5, 0.02])
    v = wall(v0, (b + x0[1]) - wy, [0, -1, 0], st, dp)
    xs = math.sin(float(np.float32(0.15) * np.float32(2.01))) * rad
    zz = abs(hc + wz)
    rr = ((x0[0] + xs) / (2 * xs)) * 0.7 * (rad * float(np.float32(0.4)))
    aa = ((x0[2] - zz / 2) / zz) * 1.6 * (2 * 3.141592654)
    xn = np.array([math.cos(aa) * rr, math.sin(aa) * rr, (wz - b) - abs(v[1] - ev) * 0.01])
    vn = np.array([v[0], v[2], -v[1]])
    x1, v1 = integrate(p, x0, v0)
    np.testing.assert_allclose(v1, vn, rtol=0, atol=2e-6)
    np.testing.assert_allclose(x1, xn + vn * DT, rtol=0, atol=2e-6)


# --------------------------------------------------- collider, accelerators
def force_side(p, x, v):
    """Velocity increment of a one-particle step (no pairs): the collider and
    accelerator terms only.  x, v are the post-integrate state when gravity
    is zero and the particle is away from the walls."""
    pos, vel = one(x, v)
    out = oracle.step_f32(p, pos, vel)
    return out["dvel"][0, :3].astype(np.float64), out["pos"][0, :3].astype(np.float64), \
        out["vel_int"][0, :3].astype(np.float64)


def test_collider_dem_closed_form():
    p = shape_params(collider=1, coll_pos=[0.02, -0.01, 0.0], coll_r=0.03, coll_spring=80.0,
                     coll_damping=0.02, coll_shear=0.1)
    dv, x1, v1 = force_side(p, [0.0, 0.0, 0.01], [0.2, -0.1, 0.3])
    c = f32(p.coll_pos)
    rel = c - x1
    dist = np.linalg.norm(rel)
    rad = float(np.float32(R) + np.float32(0.03))
    assert dist < rad
    u = -v1
    nrm = rel / dist
    tan = u - (u @ nrm) * nrm
    f = float(np.float32(80.0)) * (dist - rad) * nrm + float(np.float32(0.02)) * u + \
        float(np.float32(0.1)) * tan
    np.testing.assert_allclose(dv, f, rtol=1e-5, atol=1e-9)
    # outside the collider: nothing
    dv, _, _ = force_side(p, [0.1, 0.1, 0.1], [0.2, -0.1, 0.3])
    np.testing.assert_array_equal(dv, 0)


def test_collider_default_placement():
    """coll_pos.x == -77: placed at (wmin.x + b - 1.2 R_c, wmin.y + b + R_c, 0)
    with b = soft - r (source/SPH/Scene.cpp:100,111-112)."""
    p = shape_params(collider=1)
    b = float(np.float32(np.float32(p.dist_bnd_soft) - np.float32(R)))
    np.testing.assert_allclose(p.coll_pos[:], [-0.2 + b - 1.2 * 0.05, -0.2 + b + 0.05, 0.0], atol=1e-7)
    assert p.physics == sph.SPH_PHYS_COLLIDER


def accel(kind, pos, size, acc):
    a = sph.Accel()
    a.type = kind
    for c in range(3):
        a.pos[c], a.size[c], a.acc[c] = pos[c], size[c], acc[c]
    return a


@pytest.mark.parametrize("kind", [sph.SPH_ACC_BOX, sph.SPH_ACC_CYL_Y, sph.SPH_ACC_CYL_Y_SMOOTH])
def test_accelerators(kind):
    size = (0.03, 0.02, 0.05)
    acc = (1.0, 20.0, -3.0)
    p = shape_params(acc=[accel(kind, (0.01, 0.0, 0.0), size, acc)])
    assert p.physics == sph.SPH_PHYS_ACCEL
    inside = [0.01 + 0.01, 0.005, 0.01]
    dv, x1, _ = force_side(p, inside, [0.0, 0.0, 0.0])
    if kind == sph.SPH_ACC_CYL_Y_SMOOTH:
        r = math.hypot(x1[0] - 0.01, x1[2])
        g = 1 - r / float(np.float32(0.05))
        exp = f32(acc) * g * DT
    else:
        exp = f32(acc) * DT
    np.testing.assert_allclose(dv, exp, rtol=1e-5, atol=1e-12)
    # outside in y (all three shapes test |rel.y| < size.y)
    dv, _, _ = force_side(p, [0.01, 0.03, 0.0], [0.0, 0.0, 0.0])
    np.testing.assert_array_equal(dv, 0)
    if kind == sph.SPH_ACC_CYL_Y:
        # inside the box's corner but outside the ellipse
        dv, _, _ = force_side(p, [0.01 + 0.028, 0.0, 0.045], [0.0, 0.0, 0.0])
        np.testing.assert_array_equal(dv, 0)


# ------------------------------------------------------ rotor, height map
def dem_r(rel, u, rad, spring, damping):
    """collideSpheresR: spring (dist - rad) rel/dist + damping u inside rad."""
    dist = np.linalg.norm(rel)
    if dist >= rad:
        return np.zeros(3)
    return spring * (dist - rad) * rel / dist + damping * u


def test_pump_rotor_single_sphere():
    """One blade of one sphere (RotorSizes 0 0 1, one blade) sits at
    centre + sp (sin a, cos a, 0) with a = -rAngle + rTwist (the z = 1 row)."""
    ang, twist = 0.3, 0.2
    p = shape_params(rot_type=sph.SPH_ROTOR_PUMP_Z, rot_size=[0, 0, 1], rot_blades=1, rot_r=2.0, rot_spc=0.8,
                     rot_angle=ang, rot_twist=twist, coll_pos=[0.01, 0.0, -0.02])
    assert p.physics == sph.SPH_PHYS_ROTOR
    sp, rr = float(p.rot_spc), float(np.float32(R) + p.rot_r)
    a = -float(np.float32(ang)) + float(np.float32(twist))
    c = f32(p.coll_pos)
    sphere = c + np.array([math.sin(a) * sp, math.cos(a) * sp, 0.0])
    x = sphere + np.array([0.3, -0.5, 0.2]) * rr
    v = np.array([0.2, -0.1, 0.05])
    dv, x1, v1 = force_side(p, x, v)
    exp = dem_r(sphere - x1, -v1, rr, 80.0, float(np.float32(0.02)))
    assert np.linalg.norm(exp) > 0
    np.testing.assert_allclose(dv, exp, rtol=2e-5, atol=1e-9)
    # outside the rotor's z band nothing acts
    dv, _, _ = force_side(p, sphere + np.array([0, 0, 3 * rr]), v)
    np.testing.assert_array_equal(dv, 0)


def test_pump_rotor_y_lays_blades_in_xz():
    """Type 2 turns the same blade into the x-z plane: centre + sp (sin a, 0, cos a)."""
    ang = -0.4
    p = shape_params(rot_type=sph.SPH_ROTOR_PUMP_Y, rot_size=[0, 0, 1], rot_blades=1, rot_r=2.0, rot_spc=0.8,
                     rot_angle=ang, rot_twist=0.0, coll_pos=[0.0, 0.02, 0.0])
    sp, rr = float(p.rot_spc), float(np.float32(R) + p.rot_r)
    a = -float(np.float32(ang))
    sphere = f32(p.coll_pos) + np.array([math.sin(a) * sp, 0.0, math.cos(a) * sp])
    x = sphere + np.array([-0.4, 0.3, 0.3]) * rr
    dv, x1, v1 = force_side(p, x, np.zeros(3))
    np.testing.assert_allclose(dv, dem_r(sphere - x1, -v1, rr, 80.0, 0.02), rtol=2e-5, atol=1e-9)


def test_propeller_tip_sphere():
    """RotorSizes 0 0 0: the only sphere is the tip row (h = sy, dh = 1.2) at
    centre + (0, 0, -1.2 sp k), k = cos(-0.2 x2 + 0.6 pi), radius r + 0.2 rot_r."""
    p = shape_params(rot_type=sph.SPH_ROTOR_PROPELLER_Z, rot_size=[2, 0, 0], rot_blades=1, rot_r=2.0,
                     rot_spc=0.8, coll_pos=[0.0, 0.0, 0.0])
    sp = float(p.rot_spc)
    k = math.cos(-1.0 * 0.2 + float(np.float32(3.141592654) * np.float32(0.6)))
    sphere = np.array([0.0, 0.0, -1.2 * sp * k])
    rad = float(np.float32(R)) + float(p.rot_r) * abs(1 - 1.2)
    x = sphere + np.array([0.2, -0.3, 0.1]) * rad
    dv, x1, v1 = force_side(p, x, [0.1, 0.0, 0.0])
    exp = dem_r(sphere - x1, -v1, rad, 80.0, 0.02)
    assert np.linalg.norm(exp) > 0
    np.testing.assert_allclose(dv, exp, rtol=5e-5, atol=1e-9)


def hmap_expect(p, x, v, xz=True):
    """All 5 x 5 x 2 spheres of the sheet near x, summed in float64."""
    spc, rr = float(p.rot_spc), float(np.float32(R) + p.rot_r)
    wmin = f32(p.world_min)
    b = float(np.float32(p.dist_bnd_soft) - np.float32(R))
    wsd = f32(p.world_max) - b - (wmin + b)
    s = f32(p.shape_s)
    a_ax, l_ax = (0, 1) if xz else (1, 0)
    ia = int((x[a_ax] - wmin[a_ax]) / spc)
    iz = int((x[2] - wmin[2]) / spc)
    out = np.zeros(3)
    for j in range(-2, 3):
        for i in range(-2, 3):
            ah = ia * spc + wmin[a_ax] + i * spc
            zh = iz * spc + wmin[2] + j * spc
            ss = math.sin(s[0] * ah / wsd[a_ax] * 2 * math.pi + s[1] * math.pi / 180) * \
                math.sin(s[2] * zh / wsd[2] * 2 * math.pi + s[3] * math.pi / 180)
            if not (s[4] <= -1 or ss > s[4]):
                continue
            for k in range(2):
                sph_pos = np.zeros(3)
                sph_pos[a_ax] = ah
                sph_pos[2] = zh
                sph_pos[l_ax] = ss * float(p.h_close) + float(p.rot2_angle) + wmin[l_ax] - k * spc
                out += dem_r(sph_pos - x, -np.asarray(v), rr, 80.0, 0.02)
    return out


@pytest.mark.parametrize("plane", [sph.SPH_HMAP_XZ, sph.SPH_HMAP_YZ])
@pytest.mark.parametrize("thresh", [-1.0, 0.2])
def test_height_map_sheet(plane, thresh):
    """The sheet of spheres at h_close sin(s1 u + s2) sin(s3 w + s4) +
    rot2_angle above world_min (two layers), only where the product exceeds
    s5 (s5 <= -1: everywhere), summed over the 5 x 5 sites around the particle."""
    p = shape_params(hmap_type=plane, rot_r=2.0, rot_spc=0.7, h_close=0.01, rot2_angle=0.05,
                     shape_s=[2.0, 30.0, 3.0, 10.0, thresh, 0.9])
    assert p.physics & sph.SPH_PHYS_HMAP
    xz = plane == sph.SPH_HMAP_XZ
    lv = 1 if xz else 0
    hits = 0
    for frac in (0.013, 0.037, 0.061):
        x = np.array([0.0123, 0.0231, frac])
        x[lv] = -0.2 + 0.05 + 0.004  # just above the sheet's mean level
        dv, x1, v1 = force_side(p, x, [0.0, -0.2, 0.0])
        exp = hmap_expect(p, x1, v1, xz)
        # the collider (default place) is far away
        np.testing.assert_allclose(dv, exp, rtol=2e-4, atol=1e-8)
        hits += np.linalg.norm(exp) > 0
    assert hits > 0
