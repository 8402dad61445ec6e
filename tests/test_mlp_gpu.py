"""Learned pair-force surrogate on the device against the CPU oracle
(oracle/mlp_oracle.py).

Tolerances: the device sums in FP32 in walk order where the oracle sums in
FP64, and a hidden activation whose FP32 pre-activation lands next to a
BF16 rounding boundary can round the other way.  Probe networks whose
every sum is exact in FP32 are compared exactly; random networks within
2e-3 of the largest output (forward) and 3e-3 in relative L2 norm per
parameter block (gradient).
"""
import numpy as np
import pytest

import oracle
from oracle import mlp_oracle as mo
from pibiti_amd import scenes, sph

pytestmark = pytest.mark.gpu


def norm_for(p) -> sph.MlpNorm:
    nm = sph.MlpNorm()
    nm.inv_h = 1.0 / p.h
    nm.inv_v = 1.0
    nm.inv_rho0 = 1.0 / p.rest_density
    nm.inv_p = 1.0 / (p.stiffness * p.rest_density)
    nm.inv_nnb = 1.0 / 32.0
    nm.agg_scale = 1.0 / 32.0
    nm.a_scale = 10.0
    nm.rho_scale = 10.0
    return nm


def window_case(lattice=(10, 8, 10), before=30, window=4, seed=3):
    p, pos, vel, _ = scenes.scaled_dam_break(lattice, seed=seed)
    print("case: window", lattice, "before", before, "window", window, flush=True)
    ctx = sph.SphSystem(p, keep_dvel=True)
    ctx.reset(pos, vel)
    ctx.update(before)
    model = sph.Model(norm_for(p), seed=11)
    ctx.update(1)
    ctx.learn_begin(model)
    for _ in range(window - 1):
        ctx.update(1)
        ctx.learn_add(model)
    w = dict(pos=ctx.learn_window(sph.SPH_ARRAY_POSITION), vel=ctx.learn_window(sph.SPH_ARRAY_VELOCITY),
             rho=ctx.learn_window(sph.SPH_ARRAY_DENSITY), pres=ctx.learn_window(sph.SPH_ARRAY_PRESSURE),
             target=ctx.learn_window(sph.SPH_ARRAY_TARGET))
    assert ctx.stats()["cap_cells"] == 0
    return ctx, model, w, p


def oracle_inputs(w, p, nm):
    nd = nm.as_dict()
    pi, pj = mo.pairs_bruteforce(w["pos"], p.h)
    n = w["pos"].shape[0]
    E = mo.features(w["pos"], w["vel"], w["rho"], pi, pj, nd)
    q = mo.node_features(w["rho"], w["pres"], np.bincount(pi, minlength=n), nd)
    return E, pi, q, nd


def random_weights(seed):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, 0.5, (64, 16))
    W1[:, list(mo.PAD_COLUMNS)] = 0.0
    W2 = rng.normal(0, 0.2, (64, 64))
    b2 = rng.normal(0, 0.05, 64)
    W3 = rng.normal(0, 0.3, (3, 67))
    b3 = rng.normal(0, 0.1, 3)
    return mo.pack(W1, W2, b2, W3, b3)


def test_window_matches_plain_steps():
    """The kept window is the state the force kernel saw: its densities are
    the oracle's density of those positions, and its target is the sum of
    the steps' force increments over the window."""
    p, pos, vel, _ = scenes.scaled_dam_break((8, 6, 8), seed=4)
    with sph.SphSystem(p, keep_dvel=True) as ctx, sph.Model(norm_for(p)) as model:
        ctx.reset(pos, vel)
        ctx.update(5)
        ctx.update(1)
        ctx.learn_begin(model)
        dv = ctx.get_array(sph.SPH_ARRAY_DVEL).astype(np.float64)
        for _ in range(2):
            ctx.update(1)
            ctx.learn_add(model)
            dv += ctx.get_array(sph.SPH_ARRAY_DVEL)
        wpos = ctx.learn_window(sph.SPH_ARRAY_POSITION)
        wrho = ctx.learn_window(sph.SPH_ARRAY_DENSITY)
        tgt = ctx.learn_window(sph.SPH_ARRAY_TARGET)
    ref, _ = oracle.density_bruteforce_f64(p, wpos)
    assert np.max(np.abs(wrho - ref) / ref) < 1e-5
    np.testing.assert_allclose(tgt[:, :3], dv[:, :3] / (3 * p.time_step), rtol=1e-5,
                               atol=1e-6 * np.abs(dv).max() / p.time_step)


def test_probe_network_on_device():
    ctx, model, w, p = window_case()
    with ctx, model:
        nm = model.norm()
        model.set_weights(mo.probe_params(nm.agg_scale))
        a = ctx.learn_predict(model)
    want = mo.probe_expected(w["pos"], w["rho"], p.h, nm.as_dict())
    got = a[:, :3].astype(np.float64) / np.float32(nm.a_scale)
    assert np.array_equal(got[:, 0], want[:, 0])  # neighbour counts, exact
    np.testing.assert_allclose(got[:, 1:], want[:, 1:], rtol=2e-6, atol=1e-6)


@pytest.mark.parametrize("wseed", [1, 2])
def test_forward_matches_oracle(wseed):
    ctx, model, w, p = window_case(seed=5)
    with ctx, model:
        params = random_weights(wseed)
        model.set_weights(params)
        a = ctx.learn_predict(model)
        nm = model.norm()
    E, pi, q, nd = oracle_inputs(w, p, nm)
    ref = mo.forward(params, E, pi, w["pos"].shape[0], q, nd)["out"] * np.float32(nm.a_scale)
    err = np.abs(a[:, :3] - ref).max() / np.abs(ref).max()
    print("forward max rel err", err)
    assert err < 2e-3


def test_gradient_matches_oracle():
    ctx, model, w, p = window_case(seed=6)
    with ctx, model:
        params = random_weights(3)
        model.set_weights(params)
        loss, g = ctx.learn_gradient(model)
        nm = model.norm()
    E, pi, q, nd = oracle_inputs(w, p, nm)
    n = w["pos"].shape[0]
    target = mo.targets_from_window(w["target"], nm.a_scale)
    ref_loss, ref_g, _ = mo.loss_grad(params, E, pi, n, q, target, nd)
    assert abs(loss - ref_loss) / ref_loss < 1e-4
    blocks = dict(W1=(mo.OFF_W1, mo.OFF_W2), W2=(mo.OFF_W2, mo.OFF_B2), b2=(mo.OFF_B2, mo.OFF_W3),
                  W3=(mo.OFF_W3, mo.OFF_B3), b3=(mo.OFF_B3, mo.NUM_PARAMS))
    for name, (a, b) in blocks.items():
        rel = np.linalg.norm(g[a:b] - ref_g[a:b]) / np.linalg.norm(ref_g[a:b])
        print("grad", name, rel)
        assert rel < 3e-3, name


def test_training_reduces_loss_deterministically():
    losses = []
    for _ in range(2):
        ctx, model, w, p = window_case(seed=7)
        with ctx, model:
            l0, _ = ctx.learn_gradient(model)
            l1 = ctx.learn_train(model, 150, 2e-3)
            wts = model.weights()
        losses.append((l0, l1, wts))
    (l0, l1, w_a), (_, l1b, w_b) = losses
    print("loss", l0, "->", l1)
    assert l1 < 0.5 * l0
    assert l1 == l1b and np.array_equal(w_a, w_b)  # fixed reduction order


def test_zero_network_step_is_the_plain_integrate():
    """With every weight zero the surrogate adds nothing: one surrogate step
    is the plain integrate at factor x dt, bit for bit."""
    p, pos, vel, _ = scenes.scaled_dam_break((8, 6, 8), seed=8)
    with sph.SphSystem(p) as ctx, sph.Model(norm_for(p)) as model:
        ctx.reset(pos, vel)
        ctx.update(3)
        x0 = ctx.get_array(sph.SPH_ARRAY_POSITION)
        v0 = ctx.get_array(sph.SPH_ARRAY_VELOCITY)
        model.set_weights(np.zeros(sph.MLP_NUM_PARAMS, np.float32))
        rep = ctx.step_learned(model, 1, factor=4, tol=1e9)
        x1 = ctx.get_array(sph.SPH_ARRAY_POSITION)
        v1 = ctx.get_array(sph.SPH_ARRAY_VELOCITY)
    assert rep["coarse_steps"] == 1 and rep["fallbacks"] == 0
    q = p.copy()
    q.time_step = 4 * p.time_step
    xr, vr = oracle.integrate_f32(q, x0, v0)
    assert np.array_equal(x1, xr) and np.array_equal(v1, vr)


def test_fallback_replays_plain_steps_exactly():
    """A tolerance nothing meets undoes every surrogate step: the result is
    the plain solver's, bit for bit."""
    p, pos, vel, _ = scenes.scaled_dam_break((8, 6, 8), seed=9)
    with sph.SphSystem(p) as a, sph.SphSystem(p) as b, sph.Model(norm_for(p), seed=3) as model:
        a.reset(pos, vel)
        b.reset(pos, vel)
        a.update(2)
        b.update(2)
        rep = a.step_learned(model, 3, factor=4, tol=1e-30)
        b.update(12)
        assert rep["fallbacks"] == 3 and rep["plain_steps"] == 12 and rep["coarse_steps"] == 0
        for k in (sph.SPH_ARRAY_POSITION, sph.SPH_ARRAY_VELOCITY, sph.SPH_ARRAY_DENSITY):
            assert np.array_equal(a.get_array(k), b.get_array(k)), k
        assert abs(a.stats()["sim_time"] - b.stats()["sim_time"]) < 1e-12


def skin_case(seed=12, sigma=1.0, check=4):
    """A small dam break with fast random velocities, so that pairs cross h
    within a few surrogate steps, and a random network."""
    p, pos, vel, _ = scenes.scaled_dam_break((10, 8, 10), seed=seed, vel_sigma=sigma)
    p.tune[2] = check
    return p, pos, vel


@pytest.mark.parametrize("skin", [0.0, 0.15])
def test_learned_neighbours_and_density_match_brute_force(skin):
    """After surrogate steps (with a skin: a sort and two steps that walk its
    cell ranges) each particle's listed neighbours are exactly the pairs
    within h of its current position, and its density is the oracle's."""
    p, pos, vel = skin_case()
    print("case: skin", skin, flush=True)
    with sph.SphSystem(p) as ctx, sph.Model(norm_for(p)) as model:
        model.set_weights(random_weights(6))
        ctx.reset(pos, vel)
        ctx.update(2)
        rep = ctx.step_learned(model, 3, factor=4, tol=1e9, skin=skin)
        x = ctx.get_array(sph.SPH_ARRAY_POSITION)
        rho = ctx.get_array(sph.SPH_ARRAY_DENSITY)
        nn = ctx.get_array(sph.SPH_ARRAY_NEIGHBOURS)
        assert ctx.stats()["cap_cells"] == 0
    print("report", rep)
    assert rep["coarse_steps"] == 3 and rep["fallbacks"] == 0
    assert rep["rebuilds"] == (1 if skin else 0) and rep["early_rebuilds"] == 0
    pi, _ = mo.pairs_bruteforce(x, p.h)
    assert np.array_equal(nn, np.bincount(pi, minlength=x.shape[0]))
    ref, _ = oracle.density_bruteforce_f64(p, x)
    assert np.max(np.abs(rho - ref) / ref) < 1e-5


def test_skin_steps_follow_the_resorting_path():
    """Steps between sorts change only the order of the density and
    neighbour sums: the trajectory stays within rounding of the path that
    sorts every step, and is reproduced exactly by a second run."""
    p, pos, vel = skin_case(seed=13)
    out = {}
    for skin in (0.0, 0.15, 0.15):
        with sph.SphSystem(p) as ctx, sph.Model(norm_for(p)) as model:
            model.set_weights(random_weights(7))
            ctx.reset(pos, vel)
            ctx.update(2)
            rep = ctx.step_learned(model, 8, factor=4, tol=1e9, skin=skin)
            assert rep["coarse_steps"] == 8
            got = (ctx.get_array(sph.SPH_ARRAY_POSITION), ctx.get_array(sph.SPH_ARRAY_VELOCITY))
        if skin in out:
            assert all(np.array_equal(a, b) for a, b in zip(got, out[skin]))
        out[skin] = got
    (x0, v0), (x1, v1) = out[0.0], out[0.15]
    dv = np.abs(v1 - v0)[:, :3].max() / np.abs(v0[:, :3]).max()
    dx = np.abs(x1 - x0)[:, :3].max() / p.h
    print("skin vs sorted: max |dv| / max |v|", dv, "max |dx| / h", dx)
    assert dv < 1e-4 and dx < 1e-5


def test_skin_exceeded_rebuilds_early():
    """With a skin far thinner than one step's motion, every step past a
    sort is cut short and taken again after a new sort: no step is lost to
    the fallback, and the lists stay exact."""
    p, pos, vel = skin_case(seed=14)
    with sph.SphSystem(p) as ctx, sph.Model(norm_for(p)) as model:
        model.set_weights(random_weights(8))
        ctx.reset(pos, vel)
        ctx.update(2)
        rep = ctx.step_learned(model, 4, factor=4, tol=1e9, skin=1e-4)
        x = ctx.get_array(sph.SPH_ARRAY_POSITION)
        nn = ctx.get_array(sph.SPH_ARRAY_NEIGHBOURS)
    print("report", rep)
    assert rep["coarse_steps"] == 4 and rep["fallbacks"] == 0 and rep["plain_steps"] == 0
    # four runs, each a sort and one step; the first three also a second
    # step that is undone, and a sort replaying the kept step
    assert rep["early_rebuilds"] == 3 and rep["rebuilds"] == 7
    pi, _ = mo.pairs_bruteforce(x, p.h)
    assert np.array_equal(nn, np.bincount(pi, minlength=x.shape[0]))


def test_batch_window_gradient_matches_single_scenes():
    """A batch window's loss is the particle-weighted mean of its scenes'."""
    sc = [scenes.scaled_dam_break((8, 6, 8), seed=s) for s in (10, 11)]
    params = random_weights(4)
    single = []
    for p, pos, vel, _ in sc:
        with sph.SphSystem(p, keep_dvel=True) as c, sph.Model(norm_for(p)) as m:
            m.set_weights(params)
            c.reset(pos, vel)
            c.update(4)
            c.learn_begin(m)
            c.update(1)
            c.learn_add(m)
            single.append(c.learn_gradient(m)[0])
    with sph.SphBatch([p for p, _, _, _ in sc], keep_dvel=True) as b, \
            sph.Model(norm_for(sc[0][0])) as m:
        m.set_weights(params)
        for s, (p, pos, vel, _) in enumerate(sc):
            b.reset(s, pos, vel)
        b.update(4)
        b.learn_begin(m)
        b.update(1)
        b.learn_add(m)
        loss, _ = b.learn_gradient(m)
    ns = [int(p.num_particles) for p, _, _, _ in sc]
    want = (single[0] * ns[0] + single[1] * ns[1]) / sum(ns)
    assert abs(loss - want) / want < 1e-5
