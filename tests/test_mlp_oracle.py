"""CPU checks of the surrogate oracle (oracle/mlp_oracle.py): BF16 rounding
against torch's conversion, the analytic gradient against finite
differences, and closed-form networks against brute-force pair sums."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo

NORM = dict(inv_h=1.0 / 0.04, inv_v=1.0, inv_rho0=1.0 / 1000.0, inv_p=1.0 / 3000.0, inv_nnb=1.0 / 32.0,
            agg_scale=1.0 / 32.0, a_scale=10.0, rho_scale=10.0)


def _cloud(n=48, seed=0, h=0.04):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.0, 0.12, (n, 3))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0.0, 0.3, (n, 3))
    rho = rng.uniform(900.0, 1100.0, n).astype(np.float32)
    pres = ((rho - 1000.0) * 3.0).astype(np.float32)
    pi, pj = mo.pairs_bruteforce(pos, h)
    return pos, vel, rho, pres, pi, pj


def test_bf16_matches_torch_conversion():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-3, 1000), rng.normal(0, 1e4, 1000),
                        [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, -0.0]]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mo.bf16(x), want)
    # ties go to the even mantissa
    assert mo.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert mo.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def test_parameter_layout():
    assert mo.NUM_PARAMS == 5388
    w = np.arange(mo.NUM_PARAMS, dtype=np.float32)
    u = mo.unpack(w)
    np.testing.assert_array_equal(mo.pack(u["W1"], u["W2"], u["b2"], u["W3"], u["b3"]), w)
    assert u["W1"][2, 9] == 2 * 16 + 9 and u["W3"][1, 0] == mo.OFF_W3 + 67
    assert sorted(mo.FEATURE_COLUMNS + mo.PAD_COLUMNS + (mo.BIAS_FEAT,)) == list(range(16))


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    pos, vel, rho, pres, pi, pj = _cloud(seed=seed)
    n = pos.shape[0]
    assert pi.size > 4 * n  # a few neighbours each
    E = mo.features(pos, vel, rho, pi, pj, NORM, emulate=False)
    nnb = np.bincount(pi, minlength=n)
    q = mo.node_features(rho, pres, nnb, NORM)
    rng = np.random.default_rng(10 + seed)
    W1 = rng.normal(0, 0.6, (64, 16))
    W1[:, list(mo.PAD_COLUMNS)] = 0.0
    W2 = rng.normal(0, 0.25, (64, 64))
    b2 = rng.normal(0, 0.1, 64)
    W3 = rng.normal(0, 0.3, (3, 67))
    b3 = rng.normal(0, 0.1, 3)
    params = mo.pack(W1, W2, b2, W3, b3).astype(np.float64)
    target = rng.normal(0, 1.0, (n, 3))
    loss, grad, _ = mo.loss_grad(params, E, pi, n, q, target, NORM, emulate=False)

    def L(p):
        return mo.loss_grad(p, E, pi, n, q, target, NORM, emulate=False)[0]

    # every node-layer parameter and a sample of each pair layer
    idx = list(range(mo.OFF_W3, mo.NUM_PARAMS))
    w1_idx = np.arange(mo.OFF_W1, mo.OFF_W2).reshape(64, 16)[:, list(mo.FEATURE_COLUMNS) + [mo.BIAS_FEAT]]
    idx += list(rng.choice(w1_idx.ravel(), 40, replace=False))
    idx += list(rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 60, replace=False))
    idx += list(rng.choice(np.arange(mo.OFF_B2, mo.OFF_W3), 20, replace=False))
    eps = 1e-6
    worst = 0.0
    for k in idx:
        p1 = params.copy()
        p2 = params.copy()
        p1[k] += eps
        p2[k] -= eps
        fd = (L(p1) - L(p2)) / (2 * eps)
        worst = max(worst, abs(fd - grad[k]) / (abs(grad[k]) + 1e-4 * abs(loss) + 1e-8))
    assert worst < 1e-4, worst
    # the padding columns get exactly zero gradient
    g1 = grad[mo.OFF_W1:mo.OFF_W2].reshape(64, 16)
    assert np.all(g1[:, list(mo.PAD_COLUMNS)] == 0.0)


def test_probe_network_closed_forms():
    pos, vel, rho, pres, pi, pj = _cloud(seed=3)
    n = pos.shape[0]
    params = mo.probe_params(NORM["agg_scale"])
    E = mo.features(pos, vel, rho, pi, pj, NORM)
    q = mo.node_features(rho, pres, np.bincount(pi, minlength=n), NORM)
    out = mo.forward(params, E, pi, n, q, NORM)["out"]
    want = mo.probe_expected(pos, rho, 1.0 / NORM["inv_h"], NORM)
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-12)


def test_forward_permutation_and_translation_invariance():
    pos, vel, rho, pres, pi, pj = _cloud(seed=4)
    n = pos.shape[0]
    rng = np.random.default_rng(7)
    params = rng.normal(0, 0.3, mo.NUM_PARAMS)
    h = 1.0 / NORM["inv_h"]

    def run(p4, v4, r, pr):
        a, b = mo.pairs_bruteforce(p4, h)
        E = mo.features(p4, v4, r, a, b, NORM, emulate=False)
        q = mo.node_features(r, pr, np.bincount(a, minlength=n), NORM)
        return mo.forward(params, E, a, n, q, NORM, emulate=False)["out"]

    base = run(pos, vel, rho, pres)
    perm = rng.permutation(n)
    np.testing.assert_allclose(run(pos[perm], vel[perm], rho[perm], pres[perm]), base[perm], rtol=1e-12,
                               atol=1e-12)
    # a uniform velocity shift leaves relative velocities unchanged
    v2 = vel.copy()
    v2[:, :3] += np.float32(0.5)
    moved = run(pos, v2, rho, pres)
    assert np.max(np.abs(moved - base)) < 1e-5 * (1 + np.max(np.abs(base)))
