"""Bad inputs: a few particles with NaN or infinite coordinates in the
15,625-particle dam break.

The reference has no guard for them (its integrate and hash run whatever
arrives, source/CUDA/System.cu:165-205, Kernel_Cell.cui:5-34); the oracle
restates that arithmetic, so the step must equal the oracle's bit for bit,
NaNs in the same places:
* a NaN coordinate stays NaN and its cell is clamped into the grid (counted
  in out_of_grid); it fails the density test r2 < h2 but passes the force
  test max(minDist, |r|) < h (the max of minDist and NaN is minDist), so it
  is a force pair of every particle in its 27 cells, near or not, and turns
  their velocities NaN;
* an infinite coordinate meets a soft wall with an infinite push, and the
  push times the wall normal's zero components is NaN (inf * 0), so the
  particle's velocity and position turn NaN in every component and it then
  behaves as above.
Every other particle keeps its order and sums: keys, permutation and cell
ranges equal the oracle's exactly.
"""
import numpy as np
import pytest

import oracle
from pibiti_amd import scenes, sph

pytestmark = pytest.mark.gpu

BAD = {100: (0, np.nan), 2000: (1, np.nan), 7000: (2, np.nan), 9000: (1, np.inf), 12000: (2, -np.inf)}


def bad_scene(**kw):
    p, pos, vel, _ = scenes.cpu_dam_break_15k(**kw)
    for i, (axis, v) in BAD.items():
        pos[i, axis] = np.float32(v)
    return p, pos, vel


def same(a, b):
    """Equal with NaN in the same places (NaN payloads are not compared)."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def differing(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    bad = np.flatnonzero(~np.all((a == b) | (np.isnan(a) & np.isnan(b)), axis=1))
    return [(int(k), a[k].tolist(), b[k].tolist()) for k in bad[:6]], bad.size


@pytest.mark.parametrize("steps", [1, 3])
def test_nan_and_inf_positions_match_the_oracle(steps):
    p, pos, vel = bad_scene()
    print("case: NaN/inf positions, steps", steps, flush=True)
    with sph.SphSystem(p, keep_dvel=True) as s:
        s.reset(pos, vel)
        s.update(steps)
        g = {k: s.get_array(w) for k, w in (("pos", sph.SPH_ARRAY_POSITION), ("vel", sph.SPH_ARRAY_VELOCITY),
                                             ("dens", sph.SPH_ARRAY_DENSITY), ("pres", sph.SPH_ARRAY_PRESSURE))}
        grid = s.get_grid()
        st = s.stats()
    x, v = pos, vel
    for _ in range(steps - 1):
        o = oracle.step_f32(p, x, v)
        x, v = o["pos"], o["vel"]
    o = oracle.step_f32(p, x, v)
    for k in ("sorted_hash", "perm", "cell_start", "cell_end"):
        assert np.array_equal(grid[k], o[k]), k
    for k in ("pos", "vel", "dens", "pres"):
        assert same(g[k], o[k]), (k, differing(g[k], o[k]))
    nan_rows = np.isnan(g["pos"][:, :3]).any(axis=1)
    assert st["out_of_grid"] == o["counters"][1] == nan_rows.sum()
    assert np.all(nan_rows[list(BAD)])
    if steps == 1:
        assert nan_rows.sum() == len(BAD)
        # NaN coordinates stay where they were; infinite ones spread to all three
        for i, (axis, val) in BAD.items():
            assert np.isnan(g["pos"][i, :3]).sum() == (1 if np.isnan(val) else 3)
    # most of the fluid is untouched by the bad particles
    assert np.isfinite(g["dens"]).mean() > 0.99


def test_nan_positions_fast_mode_keep_order_and_sets():
    """FAST mode: the same integer results (keys, order, ranges, force pair
    counts) and non-finite values in the same places as the oracle."""
    p, pos, vel = bad_scene(precision=sph.SPH_PREC_FAST)
    print("case: NaN/inf positions, FAST", flush=True)
    with sph.SphSystem(p) as s:
        s.reset(pos, vel)
        s.update(1)
        grid = s.get_grid()
        nnb = s.get_array(sph.SPH_ARRAY_NEIGHBOURS)
        vel1 = s.get_array(sph.SPH_ARRAY_VELOCITY)
        dens = s.get_array(sph.SPH_ARRAY_DENSITY)
    q = p.copy()
    q.precision = sph.SPH_PREC_REFERENCE
    o = oracle.step_f32(q, pos, vel)
    for k in ("sorted_hash", "perm", "cell_start", "cell_end"):
        assert np.array_equal(grid[k], o[k]), k
    cnt = oracle.neighbour_counts_f32(q, o["pos"][o["perm"]], o["cell_start"], o["cell_end"], pred=1)
    ref = np.empty_like(cnt)
    ref[o["perm"]] = cnt
    assert np.array_equal(nnb, ref)
    assert np.array_equal(np.isfinite(dens), np.isfinite(o["dens"]))
    assert np.array_equal(np.isfinite(vel1), np.isfinite(o["vel"]))
