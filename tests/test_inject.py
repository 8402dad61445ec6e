"""Particle injection between steps: drops, rain and emitter blocks
(cSPH::Drop, source/SPH/SPH_Init.cpp:82-118; App::UpdateEmitter,
source/App/Update.cpp:62-97), against closed forms evaluated in float64."""
import math

import numpy as np
import pytest

from pibiti_amd import sph

f32 = np.float32


def xml(body: str) -> bytes:
    return f"<SPH>{body}</SPH>".encode()


def scene(attrs: str, body: str = ""):
    text = xml(f"<Scene {attrs}>{body}</Scene>")
    return sph.xml_scene(text, 0)


def inner_world(s):
    p = sph.params_from_scene(s)
    bd = float(f32(p.dist_bnd_soft) - f32(p.particle_r))
    lo = np.array(p.world_min, np.float64) + bd
    size = np.array(p.world_max, np.float64) - bd - lo
    return lo, size


@pytest.mark.parametrize("r", [0.0, 1.0, 2.5, 4.0])
def test_centre_drop_is_a_lattice_ball(r):
    """Sites (x, y, z) spc with x^2 + y^2 + z^2 <= r^2, around the centre of
    the world less the soft band and (r + 1) spacings on each side."""
    s, info = scene(f'World="0.4 0.4 0.4" Particles="5000" dropR="{r}"')
    pos, vel = sph.scene_drop(s, info)
    k = int(math.floor(r))
    sites = [(x, y, z) for z in range(-k, k + 1) for y in range(-k, k + 1) for x in range(-k, k + 1)
             if x * x + y * y + z * z <= r * r]
    assert len(pos) == len(sites)
    spc = float(info.spacing)
    db = (r + 1) * spc
    lo, size = inner_world(s)
    centre = (lo + db) + 0.5 * (size - 2 * db)
    exp = np.array(sites, np.float64) * spc + centre
    np.testing.assert_allclose(pos[:, :3], exp, rtol=0, atol=2e-7)
    assert np.all(vel == 0) and np.all(pos[:, 3] == 1)


def test_drop_stops_at_the_particle_count():
    s, info = scene('Particles="1000" dropR="3"')
    full, _ = sph.scene_drop(s, info, first=0)
    part, _ = sph.scene_drop(s, info, first=1000 - 20)
    assert len(full) > 20 and len(part) == 20
    np.testing.assert_array_equal(part, full[:20])


@pytest.mark.parametrize("bnd", [sph.SPH_BND_BOX, sph.SPH_BND_CYL_Y, sph.SPH_BND_SPHERE])
def test_random_drops_are_seeded_and_inside(bnd):
    s, info = scene(f'World="0.4 0.4 0.4" Particles="5000" dropR="2" bndType="{bnd}"')
    a, _ = sph.scene_drop(s, info, random=True, seed=5, draw=0)
    b, _ = sph.scene_drop(s, info, random=True, seed=5, draw=0)
    c, _ = sph.scene_drop(s, info, random=True, seed=5, draw=1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    lo, size = inner_world(s)
    assert np.all(a[:, :3] >= lo - 1e-6) and np.all(a[:, :3] <= lo + size + 1e-6)
    spc = float(info.spacing)
    db = 3 * spc
    centre = a[:, :3].mean(0)
    frac = (centre - (lo + db)) / (size - 2 * db)
    # high over a box or cylinder-Y floor, 3/4 up otherwise; x, z in the unit square / disc
    assert frac[1] == pytest.approx(1.0 if bnd in (sph.SPH_BND_BOX, sph.SPH_BND_CYL_Y) else 0.75, abs=1e-5)
    if bnd != sph.SPH_BND_BOX:
        assert math.hypot(frac[0] - 0.5, frac[2] - 0.5) <= 0.5 + 1e-5


def rot_t(ax_deg, ay_deg):
    """Transpose of the GL modelview Rx(ax) Ry(-ay) (float64)."""
    a, b = math.radians(ax_deg), math.radians(-ay_deg)
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    return (rx @ ry).T


@pytest.mark.parametrize("rot", [(0, 0), (90, 0), (30, 45), (-70, 200)])
def test_emitter_block(rot):
    """A size x size2 block one spacing ahead of the emitter, laid out in
    its x-y plane and moving along its z axis, both turned by the transpose
    of the emitter's modelview."""
    s, info = scene('Particles="5000"',
                    f'<Emitter size="4" size2="3" pos="0.01 -0.1 0.02" rot="{rot[0]} {rot[1]}" vel="2.5"/>')
    pos, vel = sph.scene_emit(info, 0)
    assert len(pos) == 12
    spc = float(info.spacing)
    mt = rot_t(*rot)
    grid = np.array([((x - 1.5) * spc, (y - 1.0) * spc, -spc) for y in range(3) for x in range(4)])
    exp = grid @ mt.T + np.array([0.01, -0.1, 0.02], np.float32)
    np.testing.assert_allclose(pos[:, :3], exp, rtol=0, atol=1e-6)
    np.testing.assert_allclose(vel[:, :3], np.tile(mt @ [0, 0, 2.5], (12, 1)), rtol=0, atol=2e-6)
    if rot == (90, 0):
        np.testing.assert_allclose(vel[0, :3], [0, 2.5, 0], atol=1e-6)  # straight up, as the fountains


def test_emitter_off_and_capacity():
    s, info = scene('Particles="5000"')
    pos, _ = sph.scene_emit(info, 1)
    assert len(pos) == 0
    s, info = scene('Particles="5000" EmitSize="300"')
    with pytest.raises(sph.SphError):
        sph.scene_emit(info, 0)
