"""Training the learned pair-force surrogate on plain-solver trajectories.

Everything runs on the device: a batch of dam-break scenes advances with
the plain step; every ``stride`` steps a window of ``window`` steps is kept
(the state the first step's force kernel saw, and the sum of the steps'
pair-force increments) and the model takes ``iters`` Adam steps on it.

The scenes follow the 1M dam break's physics (spacing 0.01, h = 0.02,
dt = 5e-5, k = 3, viscosity 0.1) at smaller particle counts: random block
shapes, random drop heights and random velocity noise, so that at any
moment the batch holds fluid in different phases of falling, impact and
sloshing.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import scenes, sph


def default_norm(p: sph.Params) -> sph.MlpNorm:
    """Feature units from the scene: positions in h, velocities in m/s,
    density deviations from rho0 in units of 10 %, pressure in k rho0,
    ~32 neighbours, output in 10 m/s^2."""
    nm = sph.MlpNorm()
    nm.inv_h = 1.0 / float(p.h)
    nm.inv_v = 1.0
    nm.inv_rho0 = 1.0 / float(p.rest_density)
    nm.inv_p = 1.0 / (float(p.stiffness) * float(p.rest_density))
    nm.inv_nnb = 1.0 / 32.0
    nm.agg_scale = 1.0 / 32.0
    nm.a_scale = 10.0
    nm.rho_scale = 10.0
    return nm


def training_scene(index: int, *, seed: int = 100, spacing: float = 0.01, h: float = 0.02,
                   dt: float = 5e-5, size=(12, 32), height=(8, 28)):
    """One training trajectory's scene: a block of random shape dropped from
    a random height into a box twice its footprint."""
    rng = np.random.default_rng([seed, index])
    nx = int(rng.integers(size[0], size[1] + 1))
    nz = int(rng.integers(size[0], size[1] + 1))
    ny = int(rng.integers(height[0], height[1] + 1))
    drop = float(rng.uniform(0.0, 0.6)) * ny * spacing
    sigma = float(rng.uniform(0.0, 0.1))
    b = scenes.fluid_origin(spacing)[0]
    world = (2 * nx * spacing + 2 * b, (ny * spacing + drop) * 1.6 + 2 * b, 2 * nz * spacing + 2 * b)
    p, pos, vel, info = scenes.dam_break(nx, ny, nz, spacing=spacing, h=h, world_max=world, dt=dt,
                                         viscosity=0.1, vel_sigma=sigma, seed=int(seed * 1000 + index))
    pos[:, 1] += np.float32(drop)
    info.update(drop=drop)
    return p, pos, vel, info


def train(model: sph.Model, *, scenes_count: int = 200, seed: int = 100, steps: int = 1200,
          stride: int = 8, window: int = 4, iters: int = 2, lr: float = 1e-3, lr_end: float = 1e-4,
          epochs: int = 1, size=(12, 32), height=(8, 28), log=None) -> dict:
    """Train ``model`` on ``scenes_count`` trajectories of ``steps`` plain
    steps, one window every ``stride`` steps.  Returns the loss history and
    timing."""
    if stride < window:
        raise ValueError("stride must cover the window")
    sc = [training_scene(i, seed=seed, size=size, height=height) for i in range(scenes_count)]
    total = sum(int(p.num_particles) for p, _, _, _ in sc)
    windows_per_epoch = steps // stride
    n_updates = epochs * windows_per_epoch * iters
    hist = []
    t0 = time.perf_counter()
    done = 0
    with sph.SphBatch([p for p, _, _, _ in sc], device=model.device, keep_dvel=True) as b:
        for ep in range(epochs):
            for s, (p, pos, vel, _) in enumerate(sc):
                b.reset(s, pos, vel)
            for w in range(windows_per_epoch):
                b.update(stride - window + 1)
                b.learn_begin(model)
                for _ in range(window - 1):
                    b.update(1)
                    b.learn_add(model)
                # cosine decay over all updates
                frac = done / max(1, n_updates - 1)
                cur = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * frac))
                loss = b.learn_train(model, iters, cur)
                done += iters
                hist.append(loss)
                if log and (w % 25 == 0 or w == windows_per_epoch - 1):
                    log(f"epoch {ep} window {w}/{windows_per_epoch} loss {loss:.4g} lr {cur:.3g}")
        b.synchronize()
    return dict(loss=hist, seconds=time.perf_counter() - t0, particles=total, scenes=scenes_count,
                steps=steps, stride=stride, window=window, iters=iters, epochs=epochs)


def rollout_error(params: sph.Params, pos: np.ndarray, vel: np.ndarray, model: sph.Model | None, *,
                  warm: int, coarse: int, factor: int = 4, tol: float = 1e9, skin: float = 0.0) -> dict:
    """Density error of ``coarse`` surrogate steps against factor x coarse
    plain steps from the same state (after ``warm`` plain steps):
    RMS and max over particles of |rho_surrogate - rho_plain| / rho0, and
    the same for the plain solver at factor x dt.  ``model=None`` skips the
    surrogate (plain references only)."""
    rho0 = float(params.rest_density)
    out = {}
    with sph.SphSystem(params) as ref:
        ref.reset(pos, vel)
        ref.update(warm)
        x0 = ref.get_array(sph.SPH_ARRAY_POSITION)
        v0 = ref.get_array(sph.SPH_ARRAY_VELOCITY)
        ref.update(factor * coarse)
        rho_ref = ref.get_array(sph.SPH_ARRAY_DENSITY).astype(np.float64)

    def err(rho):
        d = np.abs(rho - rho_ref) / rho0
        return dict(rms=float(np.sqrt(np.mean(d * d))), max=float(d.max()))

    # plain solver at factor x dt
    q = params.copy()
    q.time_step = factor * float(params.time_step)
    with sph.SphSystem(q) as c:
        c.reset(x0, v0)
        c.update(coarse)
        out["plain_coarse"] = err(c.get_array(sph.SPH_ARRAY_DENSITY))
    if model is not None:
        with sph.SphSystem(params) as c:
            c.reset(pos, vel)
            c.update(warm)
            rep = c.step_learned(model, coarse, factor=factor, tol=tol, skin=skin)
            out["surrogate"] = err(c.get_array(sph.SPH_ARRAY_DENSITY))
            out["report"] = rep
    return out
