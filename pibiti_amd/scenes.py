"""Seeded scene builders for the WCSPH step.

Every builder returns ``(params, pos4, vel4, info)``: derived kernel constants
(sph_params), float4 positions and velocities in original particle order, and
a dict that records every choice made (sizes, seeds, the particle radius the
reference's r-relative boundary and clamp distances are scaled by, mass).

Choices that the workloads leave open, fixed here once:
  * particle radius r = spacing / 2; soft wall band 8 r, hard clamp 1 r,
    wall stiffness 30000, wall damping 256 / 60, min_dist 1 r, min_dens 1
    (the reference defaults of source/SPH/Scene.cpp:16-31, in units of r);
  * the fluid block starts (soft - r) inside the world's min corner, the
    shift the reference applies to initMin (source/SPH/Scene.cpp:101-102),
    so no particle starts inside a soft wall band;
  * particle mass m = rho0 / (Poly6 * S) with S the self-excluded kernel sum
    over the unjittered lattice, so an interior particle starts exactly at
    the rest density (m = rho0 * spacing^3 would start the fluid at ~0.8
    rho0 because the density sum excludes the self term);
  * cell size = h, so the 27-cell search finds every pair with r < h.
"""
from __future__ import annotations

import math

import numpy as np

from . import sph


def lattice_rest_mass(spacing: float, h: float, rest_density: float, poly6: float) -> float:
    """Mass that puts an interior site of the unjittered cubic lattice at rest
    density: rho0 / (Poly6 * sum_{0 < |d| < h} (h^2 - |d|^2)^3)."""
    s = float(np.float32(spacing))
    hh = float(np.float32(h)) ** 2
    k = int(math.ceil(float(np.float32(h)) / s)) + 1
    r = np.arange(-k, k + 1) * s
    d2 = r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2
    sel = (d2 > 0) & (d2 < hh)
    total = float(np.sum((hh - d2[sel]) ** 3))
    return rest_density / (poly6 * total)


def fluid_origin(spacing: float, world_min=(0.0, 0.0, 0.0)) -> tuple:
    """Min corner of the fluid block: world_min + (soft - r) with soft = 8 r,
    r = spacing / 2."""
    b = 8.0 * (spacing / 2) - spacing / 2
    return tuple(float(w) + b for w in world_min)


def scene(n: int, h: float, spacing: float, world_min, world_max, dt: float, *,
           viscosity: float, stiffness: float = 3.0, rest_density: float = 1000.0,
           gravity=(0.0, -9.81, 0.0), eos: int = sph.SPH_EOS_LINEAR,
           precision: int = sph.SPH_PREC_REFERENCE, cfl: float = 0.0,
           max_par_in_cell: int = 16) -> sph.Scene:
    s = sph.scene_default()
    s.num_particles = n
    s.max_par_in_cell = max_par_in_cell
    s.time_step = dt
    for c in range(3):
        s.gravity[c] = gravity[c]
        s.world_min[c] = world_min[c]
        s.world_max[c] = world_max[c]
    s.global_damping = 1.0
    s.cell_size = h
    s.particle_r = spacing / 2
    s.h = h
    s.rest_density = rest_density
    s.stiffness = stiffness
    s.viscosity = viscosity
    s.min_dens = 1.0
    s.min_dist = 1.0
    s.particle_mass = 0.0
    s.dist_bnd_soft = 8.0
    s.dist_bnd_hard = 1.0
    s.bnd_stiff = 30000.0
    s.bnd_damp = 256.0
    s.bnd_damp_c = 60.0
    s.eos = eos
    s.precision = precision
    s.cfl = cfl
    poly6 = float(sph.params_from_scene(s).poly6_kern)
    s.particle_mass = lattice_rest_mass(spacing, h, rest_density, poly6)
    return s


def dam_break(nx: int, ny: int, nz: int, *, spacing: float, h: float, world_max, dt: float,
              viscosity: float, vel_sigma: float = 0.0, seed: int = 1, **kw):
    """Jittered lattice block in the corner of a box: a dam break."""
    n = nx * ny * nz
    s = scene(n, h, spacing, (0.0, 0.0, 0.0), world_max, dt, viscosity=viscosity, **kw)
    p = sph.params_from_scene(s)
    origin = fluid_origin(spacing)
    pos, vel = sph.scene_lattice(nx, ny, nz, origin, spacing, 0.1, vel_sigma,
                                 (0.0, 0.0, 0.0), seed, 0)
    info = dict(kind="dam_break", n=n, lattice=(nx, ny, nz), spacing=spacing, h=h, origin=origin,
                world_max=tuple(world_max), dt=dt, viscosity=viscosity, vel_sigma=vel_sigma,
                seed=seed, particle_r=spacing / 2, mass=float(p.particle_mass),
                grid=tuple(p.grid_size), num_cells=int(p.num_cells), key_bits=int(p.key_bits))
    return p, pos, vel, info


def cpu_dam_break_15k(seed: int = 1, **kw):
    """15,625 particles on a 25^3 lattice (spacing 0.02, a 0.5^3 block) in a
    unit box, jitter +-0.1 spacing, zero velocity, h = 0.04, rho0 = 1000, k = 3,
    viscosity 0.1, dt = 1e-4 (100 steps).  The block sits at
    fluid_origin() = (0.07, 0.07, 0.07), just inside the soft wall band."""
    return dam_break(25, 25, 25, spacing=0.02, h=0.04, world_max=(1.0, 1.0, 1.0), dt=1e-4,
                     viscosity=0.1, seed=seed, **kw)


def gpu_dam_break_1m(seed: int = 2, **kw):
    """1,048,576 particles (128x64x128 lattice, spacing 0.01, h = 0.02) in a
    2x1x2 box, jitter +-0.1 spacing, velocities N(0, 0.05) per axis, dt = 5e-5."""
    return dam_break(128, 64, 128, spacing=0.01, h=0.02, world_max=(2.0, 1.0, 2.0), dt=5e-5,
                     viscosity=0.1, vel_sigma=0.05, seed=seed, **kw)


def scaled_dam_break(lattice, *, seed: int = 3, **kw):
    """A smaller dam break with the 1M scene's resolution (spacing 0.01,
    h = 0.02, dt = 5e-5) and a box twice the block (plus the wall band) on
    every axis."""
    nx, ny, nz = lattice
    s = 0.01
    b = fluid_origin(s)[0]
    wm = (2 * nx * s + 2 * b, 2 * ny * s + 2 * b, 2 * nz * s + 2 * b)
    kw.setdefault("vel_sigma", 0.05)
    return dam_break(nx, ny, nz, spacing=s, h=0.02, world_max=wm, dt=5e-5, viscosity=0.1, seed=seed, **kw)


def colliding_blocks(edge: int = 128, *, spacing: float = 0.01, h: float = 0.02,
                     speed: float = 1.0, dt: float = 5e-5, seed: int = 4, **kw):
    """Two edge^3 lattice blocks moving at +-speed along x toward each other in
    a (4 x 2 x 2) * (edge / 128) box, no gravity."""
    scale = edge / 128.0
    world = (4.0 * scale, 2.0 * scale, 2.0 * scale)
    ext = edge * spacing
    gap = 0.24 * scale
    y0 = (world[1] - ext) / 2
    z0 = (world[2] - ext) / 2
    xa = world[0] / 2 - gap / 2 - ext
    xb = world[0] / 2 + gap / 2
    n = 2 * edge ** 3
    kw.setdefault("gravity", (0.0, 0.0, 0.0))
    s = scene(n, h, spacing, (0.0, 0.0, 0.0), world, dt, viscosity=0.1, **kw)
    p = sph.params_from_scene(s)
    pa, va = sph.scene_lattice(edge, edge, edge, (xa, y0, z0), spacing, 0.1, 0.0,
                               (speed, 0.0, 0.0), seed, 0)
    pb, vb = sph.scene_lattice(edge, edge, edge, (xb, y0, z0), spacing, 0.1, 0.0,
                               (-speed, 0.0, 0.0), seed, 1)
    pos = np.concatenate([pa, pb])
    vel = np.concatenate([va, vb])
    info = dict(kind="colliding_blocks", n=n, edge=edge, spacing=spacing, h=h, world=world,
                speed=speed, dt=dt, seed=seed, particle_r=spacing / 2,
                mass=float(p.particle_mass), grid=tuple(p.grid_size), num_cells=int(p.num_cells))
    return p, pos, vel, info


def batch_counts(nscenes: int = 64, lo: int = 2000, hi: int = 200000, seed: int = 5) -> list[int]:
    """Particle counts log-uniform in [lo, hi] from the seed."""
    rng = np.random.default_rng(seed)
    u = rng.random(nscenes)
    return [int(round(math.exp(math.log(lo) + x * (math.log(hi) - math.log(lo))))) for x in u]


def batch_lattice(count: int) -> tuple[int, int, int]:
    """Lattice dims (nx, ny, nz) with nx = nz and nx*ny*nz close to count."""
    nx = max(2, int(round(count ** (1.0 / 3.0))))
    ny = max(2, int(round(count / (nx * nx))))
    return nx, ny, nx


def batch_scene(count: int, index: int, *, seed: int = 5, **kw):
    """One scene of the batch: a dam break at the 15k scene's resolution
    (spacing 0.02, h = 0.04, dt = 1e-4) whose box is twice the block per axis
    (the 15k scene's fill fraction of 1/8)."""
    nx, ny, nz = batch_lattice(count)
    s = 0.02
    b = fluid_origin(s)[0]
    wm = (2 * nx * s + 2 * b, 2 * ny * s + 2 * b, 2 * nz * s + 2 * b)
    return dam_break(nx, ny, nz, spacing=s, h=0.04, world_max=wm, dt=1e-4, viscosity=0.1,
                     seed=seed * 1000 + index, **kw)
