"""CPU oracle of the learned pair-force surrogate - TEST INFRASTRUCTURE ONLY.

Only tests/ and bench.py's checks may import this module, as the checker;
the product library never calls it.

The surrogate has no counterpart in the reference (it is an extension of
the WCSPH step, whose pair set it reuses: the density kernel's pairs
r2 < h2, j != i, source/CUDA/Kernel_Cell.cui:160-170).  This module
restates its mathematics in float64 numpy:

  features(...)      16 edge features per pair, computed in float32 with the
                     kernel's operation order, then rounded to bfloat16
  forward(...)       pair network, neighbour sum, node layer
  loss_grad(...)     loss and analytic gradient by reverse-mode chain rule

With ``emulate=True`` the BF16 rounding points of the device kernels are
reproduced (weights, features, both layers' activations, b2 as a BF16
high/low pair, and the backward products' inputs); everything else is
float64, so the device result differs only by FP32 accumulation order (and
rare BF16 rounding-boundary flips that order can cause).

Parity pinning status (tests/test_mlp_oracle.py, tests/test_mlp_gpu.py):
  loss_grad ... central finite differences of the float64 loss (emulate=False)
  forward ..... closed forms: a network that counts neighbours and one that
                sums |x_i - x_j|/h, checked against brute-force pair sums;
                both also run on the device
  features .... the same closed forms (per-feature probes)
"""
from __future__ import annotations

import numpy as np

HID, FEAT, BIAS_FEAT, OUT, NODE = 64, 16, 5, 3, 3
# W1 columns that carry a feature (the rest of 0..15 besides BIAS_FEAT are padding)
FEATURE_COLUMNS = (0, 1, 2, 3, 4, 8, 9, 10, 11)
PAD_COLUMNS = (6, 7, 12, 13, 14, 15)
W3_COLS = HID + NODE
OFF_W1 = 0
OFF_W2 = OFF_W1 + HID * FEAT
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + OUT * W3_COLS
NUM_PARAMS = OFF_B3 + OUT


def bf16(x) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) and return as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray) -> dict:
    p = np.asarray(params, np.float64)
    return dict(
        W1=p[OFF_W1:OFF_W2].reshape(HID, FEAT),
        W2=p[OFF_W2:OFF_B2].reshape(HID, HID),
        b2=p[OFF_B2:OFF_W3].copy(),
        W3=p[OFF_W3:OFF_B3].reshape(OUT, W3_COLS),
        b3=p[OFF_B3:NUM_PARAMS].copy(),
    )


def pack(W1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), np.ravel(W2), np.ravel(b2), np.ravel(W3),
                           np.ravel(b3)]).astype(np.float32)


def pairs_bruteforce(pos: np.ndarray, h: float) -> tuple[np.ndarray, np.ndarray]:
    """All ordered pairs (i, j), j != i, with r2 < h2 evaluated in float32 as
    the device does (((dx*dx + dy*dy) + dz*dz), no contraction)."""
    x = np.asarray(pos, np.float32)[:, :3]
    h32 = np.float32(h)
    h2 = np.float32(h32 * h32)
    ii, jj = [], []
    n = x.shape[0]
    for i in range(n):
        d = (x[i][None, :] - x).astype(np.float32)
        r2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32) + d[:, 2] * d[:, 2]).astype(np.float32)
        sel = np.nonzero((r2 < h2) & (np.arange(n) != i))[0]
        ii.append(np.full(sel.size, i, np.int64))
        jj.append(sel.astype(np.int64))
    return np.concatenate(ii), np.concatenate(jj)


def features(pos, vel, rho, pi, pj, norm: dict, emulate: bool = True) -> np.ndarray:
    """Edge features (pairs x 16) in the kernel's float32 operation order."""
    f32 = np.float32
    x = np.asarray(pos, f32)[:, :3]
    v = np.asarray(vel, f32)[:, :3]
    r = np.asarray(rho, f32)
    inv_h, inv_v, inv_rho0 = f32(norm["inv_h"]), f32(norm["inv_v"]), f32(norm["inv_rho0"])
    rs = f32(norm.get("rho_scale", 1.0) or 1.0)
    d = (x[pi] - x[pj]).astype(f32)
    dist = np.sqrt(((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(f32) + d[:, 2] * d[:, 2]).astype(f32))
    dv = (v[pj] - v[pi]).astype(f32)
    F = np.zeros((pi.size, FEAT), np.float32)
    F[:, 0:3] = d * inv_h
    F[:, 3] = dist * inv_h
    F[:, 4] = ((r[pi] * inv_rho0).astype(f32) - f32(1.0)) * rs
    F[:, BIAS_FEAT] = 1.0
    F[:, 8:11] = dv * inv_v
    F[:, 11] = ((r[pj] * inv_rho0).astype(f32) - f32(1.0)) * rs
    return bf16(F) if emulate else F.astype(np.float64)


def node_features(rho, pres, nnb, norm: dict) -> np.ndarray:
    f32 = np.float32
    q = np.zeros((np.asarray(rho).size, NODE), np.float64)
    rs = f32(norm.get("rho_scale", 1.0) or 1.0)
    q[:, 0] = ((np.asarray(rho, f32) * f32(norm["inv_rho0"])).astype(f32) - f32(1.0)) * rs
    q[:, 1] = np.asarray(pres, f32) * f32(norm["inv_p"])
    q[:, 2] = np.asarray(nnb, f32) * f32(norm["inv_nnb"])
    return q


def forward(params, E, pi, n, q, norm: dict, emulate: bool = True) -> dict:
    """Pair network, neighbour sum and node layer.  Returns the network
    output (a / a_scale) and the intermediates the gradient needs."""
    w = unpack(params)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    W1, W2 = rnd(w["W1"]), rnd(w["W2"])
    b2 = w["b2"]
    if emulate:
        # the kernel adds b2 as two BF16 parts, hi = bf16(b2), lo = bf16(b2 - hi)
        b32 = np.asarray(b2, np.float32)
        hi = bf16(b32)
        lo = bf16((b32 - hi.astype(np.float32)).astype(np.float32))
        b2 = hi + lo
    pre1 = E @ W1.T
    h1 = rnd(np.maximum(pre1, 0.0))
    pre2 = h1 @ W2.T + b2[None, :]
    # pair outputs enter the neighbour sum as BF16 (FP32 accumulation)
    m = rnd(np.maximum(pre2, 0.0))
    H = np.zeros((n, HID))
    np.add.at(H, pi, m)
    agg = float(np.float32(norm["agg_scale"]))
    z = np.concatenate([agg * H, q], axis=1)
    out = z @ w["W3"].T + w["b3"][None, :]
    return dict(out=out, H=H, z=z, pre1=pre1, h1=h1, pre2=pre2, W1=W1, W2=W2)


def loss_grad(params, E, pi, n, q, target, norm: dict, emulate: bool = True):
    """Loss mean_i |out_i - target_i|^2 (target in a_scale units) and its
    gradient with respect to the parameter vector."""
    w = unpack(params)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    f = forward(params, E, pi, n, q, norm, emulate)
    diff = f["out"] - target
    loss = float(np.sum(diff * diff) / n)
    d_out = 2.0 * diff / n
    gW3 = d_out.T @ f["z"]
    gb3 = d_out.sum(axis=0)
    agg = float(np.float32(norm["agg_scale"]))
    dH = agg * (d_out @ w["W3"][:, :HID])
    dpre2 = dH[pi] * (f["pre2"] > 0.0)
    gb2 = dpre2.sum(axis=0)
    dp2 = rnd(dpre2)
    gW2 = dp2.T @ f["h1"]
    dh1 = dp2 @ f["W2"]
    dp1 = rnd(dh1 * (f["pre1"] > 0.0))
    gW1 = dp1.T @ E
    grad = np.concatenate([gW1.ravel(), gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad, f


def targets_from_window(target_accel: np.ndarray, a_scale: float) -> np.ndarray:
    """Window-mean pair-force acceleration in network output units."""
    return np.asarray(target_accel, np.float64)[:, :3] / float(np.float32(a_scale))


def probe_params(agg: float) -> np.ndarray:
    """Three independent paths through the network (hidden unit pairs chosen
    in both halves of the layer):
      a_x / a_scale = number of neighbours
      a_y / a_scale = sum_j bf16(|x_i - x_j| / h)
      a_z / a_scale = sum_j bf16(max(0, 1 + bf16(s (rho_j / rho0 - 1))))"""
    W1 = np.zeros((64, 16))
    W2 = np.zeros((64, 64))
    W3 = np.zeros((3, 67))
    W1[0, BIAS_FEAT] = 1.0
    W2[5, 0] = 1.0
    W3[0, 5] = 1.0 / agg
    W1[37, 3] = 1.0
    W2[50, 37] = 1.0
    W3[1, 50] = 1.0 / agg
    W1[20, 11] = 1.0
    W1[20, BIAS_FEAT] = 1.0
    W2[60, 20] = 1.0
    W3[2, 60] = 1.0 / agg
    return pack(W1, W2, np.zeros(64), W3, np.zeros(3))


def probe_expected(pos, rho, h: float, norm: dict) -> np.ndarray:
    """Brute-force pair sums the probe network must reproduce."""
    pi, pj = pairs_bruteforce(pos, h)
    n = pos.shape[0]
    f32 = np.float32
    x = np.asarray(pos, f32)[:, :3]
    d = (x[pi] - x[pj]).astype(f32)
    dist = np.sqrt(((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(f32) + d[:, 2] * d[:, 2]).astype(f32))
    fr = bf16((dist * f32(norm["inv_h"])).astype(f32))
    rs = f32(norm.get("rho_scale", 1.0) or 1.0)
    frho = bf16(np.maximum(0.0, 1.0 + bf16(((np.asarray(rho, f32)[pj] * f32(norm["inv_rho0"])).astype(f32)
                                             - f32(1.0)) * rs)))
    want = np.zeros((n, 3))
    want[:, 0] = np.bincount(pi, minlength=n)
    want[:, 1] = np.bincount(pi, weights=fr, minlength=n)
    want[:, 2] = np.bincount(pi, weights=frho, minlength=n)
    return want
