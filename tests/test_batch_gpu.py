"""Batched scenes, graph replay, persistent gathers and the autotuner.

Every launch variant computes the same arithmetic in the same per-particle
order, so each check below is bit-exact: a scene inside a batch against the
same scene run alone (and against the FP32 oracle), graph replays against
eager steps, persistent-grid gathers against one-thread-per-particle ones.
"""
import numpy as np
import pytest

import oracle
from pibiti_amd import scenes, sph

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "dens", "pres")


def read_scene(b, s):
    return {k: b.get_array(s, i) for i, k in enumerate(FIELDS)}


def read_ctx(c):
    return {k: c.get_array(i) for i, k in enumerate(FIELDS)}


def small_batch(counts, seed=7, **kw):
    out = []
    for i, n in enumerate(counts):
        p, pos, vel, _ = scenes.batch_scene(n, i, seed=seed, **kw)
        out.append((p, pos, vel))
    return out


def test_batch_scenes_equal_individual_runs_and_oracle():
    counts = [2100, 5000, 800, 12000, 3300, 64]
    sc = small_batch(counts)
    print("case: batch of", len(sc), "scenes,", sum(int(p.num_particles) for p, _, _ in sc), "particles",
          flush=True)
    with sph.SphBatch([p for p, _, _ in sc]) as b:
        for s, (p, x, v) in enumerate(sc):
            b.reset(s, x, v)
        b.update(12)
        got = [read_scene(b, s) for s in range(len(sc))]
        info = b.info()
    assert info["particles"] == sum(int(p.num_particles) for p, _, _ in sc)
    for s, (p, x, v) in enumerate(sc):
        with sph.SphSystem(p) as c:
            c.reset(x, v)
            c.update(12)
            alone = read_ctx(c)
        for k in FIELDS:
            assert np.array_equal(got[s][k], alone[k]), (s, k)
    # and one scene against the oracle directly
    p, x, v = sc[1]
    o = oracle.run_f32(p, x, v, 12)
    for k in FIELDS:
        assert np.array_equal(got[1][k], o[k]), k


def test_batch_graph_replay_equals_eager_steps():
    sc = small_batch([3000, 700, 9000], seed=3)
    params = [p for p, _, _ in sc]
    with sph.SphBatch(params) as eager, sph.SphBatch(params) as graphed:
        for s, (p, x, v) in enumerate(sc):
            eager.reset(s, x, v)
            graphed.reset(s, x, v)
        graphed.capture(6)
        for rep in range(3):
            graphed.replay(1)
            eager.update(6)
            # eager work and synchronisations between replays
            for s in range(len(sc)):
                a, g = read_scene(eager, s), read_scene(graphed, s)
                for k in FIELDS:
                    assert np.array_equal(a[k], g[k]), (rep, s, k)
        st = graphed.stats(0)
        assert st["steps"] == 18


def test_single_scene_graph_replay_and_persistent_grid():
    p, x, v, _ = scenes.scaled_dam_break((24, 12, 24), seed=5)
    with sph.SphSystem(p) as c:
        c.reset(x, v)
        c.update(8)
        ref = read_ctx(c)
    q = p.copy()
    q.tune[1] = 512  # persistent gathers: 512 blocks pulling tiles
    with sph.SphSystem(q) as c:
        c.reset(x, v)
        c.capture(4)
        c.replay(2)
        got = read_ctx(c)
    for k in FIELDS:
        assert np.array_equal(ref[k], got[k]), k


@pytest.mark.parametrize("tune", [(3, 0), (2, 0), (0, 0), (1, 0), (24, 0), (32, 256), (16, 1024), (16, 0, 0, 1),
                                  (24, 0, 0, 2), (16, 0, 0, 3), (48, 0, 0, 1)])
def test_launch_variants_identical(tune):
    sc = small_batch([4000, 9000], seed=11)
    base = None
    for t in [(16, 0), tune]:
        with sph.SphBatch([p for p, _, _ in sc]) as b:
            b.set_tune(t)
            for s, (p, x, v) in enumerate(sc):
                b.reset(s, x, v)
            b.update(5)
            got = [read_scene(b, s) for s in range(len(sc))]
        if base is None:
            base = got
        else:
            for s in range(len(sc)):
                for k in FIELDS:
                    assert np.array_equal(base[s][k], got[s][k]), (t, s, k)


def test_end_time_freezes_each_scene():
    """t_end = 10 ms at dt = 1e-4: after 130 steps each scene equals 100 steps alone."""
    sc = small_batch([1500, 4000], seed=13)
    params = []
    for p, _, _ in sc:
        q = p.copy()
        q.t_end = 0.01
        params.append(q)
    with sph.SphBatch(params) as b:
        for s, (p, x, v) in enumerate(sc):
            b.reset(s, x, v)
        b.update(130)
        got = [read_scene(b, s) for s in range(len(sc))]
        clocks = [b.stats(s)["sim_time"] for s in range(len(sc))]
    for s, (p, x, v) in enumerate(sc):
        assert abs(clocks[s] - 0.01) < 1e-6
        o = oracle.run_f32(p, x, v, 100)
        for k in ("pos", "vel"):
            # after the last full step the velocity update uses dt = 0
            assert np.array_equal(got[s][k], o[k]), (s, k)


def test_cfl_time_step_per_scene():
    sc = small_batch([2000, 2000], seed=17)
    params = []
    for i, (p, _, _) in enumerate(sc):
        q = p.copy()
        q.cfl = 0.002
        params.append(q)
    vfast = sc[1][2].copy()
    vfast[:, 0] += np.float32(3.0)
    with sph.SphBatch(params) as b:
        b.reset(0, sc[0][1], sc[0][2])
        b.reset(1, sc[1][1], vfast)
        d0, d1 = b.stats(0)["dt_cfl"], b.stats(1)["dt_cfl"]
    assert d1 < d0 <= float(params[0].time_step)
    vmax = np.sqrt((vfast[:, :3].astype(np.float64) ** 2).sum(1)).max()
    expect = float(params[1].cfl) * float(params[1].h) / (float(params[1].sound_speed) + vmax)
    assert abs(d1 - expect) <= 1e-15


def test_autotune_restores_state_and_caches(tmp_path):
    sc = small_batch([6000, 20000, 3000], seed=19)
    cache = tmp_path / "tune.jsonl"
    with sph.SphBatch([p for p, _, _ in sc]) as b:
        for s, (p, x, v) in enumerate(sc):
            b.reset(s, x, v)
        b.update(2)
        before = [read_scene(b, s)["pos"] for s in range(len(sc))]
        tune, ms = b.autotune(str(cache))
        after = [read_scene(b, s)["pos"] for s in range(len(sc))]
        assert tune[0] in (3, 2, 16, 24, 32) and ms > 0
        for s in range(len(sc)):
            assert np.array_equal(before[s], after[s])
        assert b.stats(0)["steps"] == 2
        tune2, ms2 = b.autotune(str(cache))  # served from the cache
        assert tune2 == tune and abs(ms2 - ms) < 1e-3
    assert cache.read_text().count("\n") == 1
