"""The batched workload at full size: 64 independent dam breaks with particle
counts log-uniform in [2k, 200k] (seed 5; 2.8M particles in all), boxes
scaled to each count, advanced as one particle array.  The autotuner picks
the launch shape on the live state (and restores it), 10 steps are captured
in a hipGraph, and 10 replays advance every scene by 10 simulated ms.

Every scene at full size: its slots are a permutation in (cell key, index)
order of its own positions, its cell ranges cover exactly its particles,
its FP64 mass is exact and its clock reads 0.01 s.  The smallest, the
median and the largest scene equal the oracle's 100 steps bit for bit.
"""
import numpy as np
import pytest

import oracle
from pibiti_amd import scenes, sph

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def test_64_scene_batch_autotuned_graph_replays():
    counts = scenes.batch_counts(64)
    sc = [scenes.batch_scene(c, i) for i, c in enumerate(counts)]
    params = [q for q, _, _, _ in sc]
    print("case: 64-scene batch,", sum(counts), "particles", flush=True)
    with sph.SphBatch(params) as b:
        for k, (_, x, v, _) in enumerate(sc):
            b.reset(k, x, v)
        tune, ms = b.autotune(None)
        print("tune", tune, "ms per step", ms, flush=True)
        b.capture(10)
        b.replay(10)
        b.synchronize()
        for k, (q, _, _, _) in enumerate(sc):
            n = int(q.num_particles)
            x = b.get_array(k, sph.SPH_ARRAY_POSITION)
            g = b.get_grid(k)
            st = b.stats(k)
            h, oob = oracle.hash_f32(q, x)
            order = np.lexsort((np.arange(n), h)).astype(np.uint32)
            assert np.array_equal(g["perm"], order), f"scene {k}: order"
            assert np.array_equal(g["sorted_hash"], h[order]), f"scene {k}: keys"
            cs, ce = g["cell_start"], g["cell_end"]
            occ = cs != EMPTY
            cnt = np.bincount(h, minlength=len(cs))
            assert np.array_equal(np.flatnonzero(occ), np.flatnonzero(cnt)), f"scene {k}: occupied cells"
            assert int(np.sum(ce[occ].astype(np.int64) - cs[occ])) == n, f"scene {k}: sum of ranges"
            assert np.array_equal(cs[occ], np.searchsorted(h[order], np.flatnonzero(cnt))), f"scene {k}: starts"
            assert st["mass"] == n * float(np.float32(q.particle_mass)), f"scene {k}: mass"
            assert abs(st["sim_time"] - 0.01) < 1e-8, f"scene {k}: clock {st['sim_time']}"
            assert st["sort_errors"] == 0
        idx = np.argsort(counts, kind="stable")
        for k in (int(idx[0]), int(idx[len(idx) // 2]), int(idx[-1])):
            q, x0, v0, _ = sc[k]
            print("case: oracle 100 steps of scene", k, counts[k], "particles", flush=True)
            o = oracle.run_f32(q, x0, v0, 100)
            for name, which in (("pos", sph.SPH_ARRAY_POSITION), ("vel", sph.SPH_ARRAY_VELOCITY),
                                ("dens", sph.SPH_ARRAY_DENSITY)):
                g = b.get_array(k, which)
                bad = np.flatnonzero(np.any((g != o[name]).reshape(len(g), -1), axis=1))
                assert bad.size == 0, f"scene {k} {name}: {bad.size} particles differ"
