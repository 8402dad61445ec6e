"""Derived constants (Scene::_UpdatePar/_UpdateGrid, source/SPH/Scene.cpp:76-113)
against their closed forms, the force-predicate threshold against brute
force, and the seeded lattice initialiser."""
import math

import numpy as np
import pytest

from pibiti_amd import scenes, sph

PI_F = np.float32(3.141592654)


def ulp_close(a, b, ulps=2):
    a = np.float32(a)
    b = np.float32(b)
    return abs(int(a.view(np.int32)) - int(b.view(np.int32))) <= ulps


@pytest.mark.parametrize("h", [0.01, 0.02, 0.04, 0.0375])
def test_kernel_constants_closed_form(h):
    s = sph.scene_default()
    s.h = h
    s.cell_size = h
    p = sph.params_from_scene(s)
    hd = float(np.float32(h))
    pi = float(PI_F)
    assert ulp_close(p.poly6_kern, 315.0 / (64.0 * pi * hd ** 9))
    assert ulp_close(p.spiky_kern, 22.5 / (pi * hd ** 6))
    assert ulp_close(p.lap_kern, 45.0 / (pi * hd ** 6))
    assert p.h2 == np.float32(np.float32(h) * np.float32(h))
    # spiky gradient magnitude coefficient is half the Laplacian's
    assert ulp_close(2 * p.spiky_kern, p.lap_kern)


def test_scaled_quantities():
    s = sph.scene_default()
    p = sph.params_from_scene(s)
    r = np.float32(s.particle_r)
    assert p.min_dist == np.float32(np.float32(s.min_dist) * r)
    assert p.dist_bnd_soft == np.float32(8 * r) and p.dist_bnd_hard == np.float32(r)
    assert ulp_close(p.min_dens, 1.0 / (1000.0 ** 2))
    # reference mass: rho0 * 4/3 * pi * r^3
    assert ulp_close(p.particle_mass, 1000.0 * 4.0 / 3.0 * math.pi * float(r) ** 3, ulps=4)
    s.particle_mass = 0.125
    assert sph.params_from_scene(s).particle_mass == np.float32(0.125)


def test_grid_shape_and_key_bits():
    s = sph.scene_default()
    p = sph.params_from_scene(s)
    # world 0.4 x 0.5 x 0.4, cell 0.008 -> 50 x 63 x 50 (ceil of 62.5)
    assert tuple(p.grid_size) == (50, 63, 50)
    assert p.num_cells == 50 * 63 * 50 and p.grid_size_yx == 50 * 63
    assert (1 << p.key_bits) >= p.num_cells > (1 << (p.key_bits - 1))
    p1, *_ = scenes.cpu_dam_break_15k()
    assert tuple(p1.grid_size) == (25, 25, 25) and p1.key_bits == 14
    p2, *_ = scenes.gpu_dam_break_1m()
    assert tuple(p2.grid_size) == (100, 50, 100) and p2.key_bits == 19


@pytest.mark.parametrize("h,min_dist", [(0.04, 0.01), (0.02, 0.005), (0.01, 0.004), (0.0375, 0.0),
                                        (0.03, 0.029999)])
def test_force_threshold_is_exact(h, min_dist):
    """force_r2_max is the largest float r2 with max(min_dist, sqrtf(r2)) < h."""
    s = sph.scene_default()
    s.h = h
    s.cell_size = h
    s.particle_r = 1.0
    s.min_dist = min_dist
    p = sph.params_from_scene(s)
    t = np.float32(p.force_r2_max)
    hf = np.float32(h)
    md = np.float32(p.min_dist)
    pred = lambda r2: np.maximum(md, np.sqrt(np.float32(r2))) < hf  # noqa: E731
    assert pred(t)
    assert not pred(np.nextafter(t, np.float32(np.inf)))
    # brute force over a window of floats around h*h
    hh = np.float32(hf * hf)
    lo = hh.view(np.int32) - 4000
    cand = np.arange(lo, lo + 8000, dtype=np.int32).view(np.float32)
    ok = np.maximum(md, np.sqrt(cand)) < hf
    assert np.array_equal(ok, cand <= t)


def test_force_threshold_when_min_dist_reaches_h():
    s = sph.scene_default()
    s.particle_r = 1.0
    s.min_dist = float(s.h)  # min_dist == h: nothing interacts
    assert sph.params_from_scene(s).force_r2_max == -1.0


def test_lattice_layout_and_jitter():
    nx, ny, nz, s = 5, 4, 3, 0.02
    pos, vel = sph.scene_lattice(nx, ny, nz, (0.1, 0.2, 0.3), s, 0.1, 0.0, (0, 0, 0), 7, 0)
    assert pos.shape == (60, 4) and np.all(pos[:, 3] == 1) and np.all(vel == 0)
    idx = np.arange(60)
    ix, iz, iy = idx % nx, (idx // nx) % nz, idx // (nx * nz)  # x fastest, then z, then y
    site = np.stack([0.1 + (ix + 0.5) * s, 0.2 + (iy + 0.5) * s, 0.3 + (iz + 0.5) * s], 1)
    dev = np.abs(pos[:, :3] - site)
    assert dev.max() <= 0.1 * s * (1 + 1e-5) + 1e-7
    assert dev.max() > 0.05 * s  # jitter is actually applied


def test_lattice_seeded_and_velocity_distribution():
    a = sph.scene_lattice(20, 20, 20, (0, 0, 0), 0.01, 0.1, 0.05, (0, 0, 0), 11, 0)
    b = sph.scene_lattice(20, 20, 20, (0, 0, 0), 0.01, 0.1, 0.05, (0, 0, 0), 11, 0)
    c = sph.scene_lattice(20, 20, 20, (0, 0, 0), 0.01, 0.1, 0.05, (0, 0, 0), 12, 0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    v = a[1][:, :3].astype(np.float64)
    assert abs(v.mean()) < 0.003
    assert abs(v.std() - 0.05) < 0.002
    # the three axes are not copies of each other
    assert abs(np.corrcoef(v[:, 0], v[:, 1])[0, 1]) < 0.05


def test_batch_counts_log_uniform():
    counts = scenes.batch_counts(64)
    assert len(counts) == 64 and min(counts) >= 2000 and max(counts) <= 200000
    assert counts == scenes.batch_counts(64)
    logs = np.log(counts)
    assert abs(logs.mean() - (np.log(2000) + np.log(200000)) / 2) < 0.6
