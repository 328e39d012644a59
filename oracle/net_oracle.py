"""CPU oracle of the learned step corrector (numpy).

TEST INFRASTRUCTURE ONLY: the checker the device perceptron is compared
against. Only tests/ (and bench.py's report) import it; the product path
never does.

The corrector has no counterpart in the reference: the reference
integrates with plain RK4 only (orbital.cpp:107-118). What this module
restates is the library's own definition of the perceptron, in two forms:

* ``forward_exact`` / ``loss_and_grad_exact``: FP64, no rounding. The
  analytic backward pass is pinned by central finite differences of the
  FP64 loss (tests/test_net_oracle.py), so a dropped term, a transposed
  weight or a wrong activation derivative fails there.
* ``forward_device`` / ``grad_device``: the same network with BF16 rounding
  at every point where the device rounds (operands of the matrix-core
  products; the input and output layers split operands into BF16 high and
  low parts and add hi*hi + hi*lo + lo*hi), FP64 accumulation. The device
  accumulates in FP32 in another order and evaluates the activation in FP32
  (one fused multiply-add, a clamp and a product), so comparisons carry a
  tolerance stated in the tests.

The activation is the hard swish z * clamp(z / 6 + 1 / 2, 0, 1) (Howard et
al., "Searching for MobileNetV3", 2019), whose derivative is 0 below -3, 1
above 3 and z / 3 + 1 / 2 between.

Parameter vector layout (35,206 floats): W1[128][10], b1[128], W2[128][128],
b2[128], W3[128][128], b3[128], W4[6][128], b4[6].
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 10, 128, 6
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NH
OFF_W4 = OFF_B3 + NH
OFF_B4 = OFF_W4 + NOUT * NH
NPARAM = OFF_B4 + NOUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NIN), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(NH, NH), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(NH, NH), p[OFF_B3:OFF_W4],
            p[OFF_W4:OFF_B4].reshape(NOUT, NH), p[OFF_B4:NPARAM])


def pack(W1, b1, W2, b2, W3, b3, W4, b4):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3, W4.ravel(), b4])


def init_params(seed=0):
    """Uniform +-1/sqrt(fan_in) initialisation (numpy RNG; any seed will do
    for the checks)."""
    rng = np.random.default_rng(seed)
    parts = []
    for nin, nout in ((NIN, NH), (NH, NH), (NH, NH), (NH, NOUT)):
        b = 1.0 / np.sqrt(nin)
        parts.append(rng.uniform(-b, b, (nout, nin)))
        parts.append(rng.uniform(-b, b, nout))
    return pack(*parts).astype(np.float32)


def act(z):
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def dact(z):
    return np.where(z < -3.0, 0.0, np.where(z > 3.0, 1.0, z / 3.0 + 0.5))


# ---------------------------------------------------------------------------
# FP64, no rounding
# ---------------------------------------------------------------------------

def forward_exact(p, x):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1.T + b1
    h1 = act(z1)
    z2 = h1 @ W2.T + b2
    h2 = act(z2)
    z3 = h2 @ W3.T + b3
    h3 = act(z3)
    return h3 @ W4.T + b4, (x, z1, h1, z2, h2, z3, h3)


def loss_and_grad_exact(p, x, y):
    """Loss mean_b (1/6) sum_k (o_bk - y_bk)^2 and its gradient (FP64)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    o, (x, z1, h1, z2, h2, z3, h3) = forward_exact(p, x)
    B = x.shape[0]
    d = o - np.asarray(y, dtype=np.float64)
    loss = float((d * d).sum() / (6.0 * B))
    g4 = 2.0 * d / (6.0 * B)
    g3 = (g4 @ W4) * dact(z3)
    g2 = (g3 @ W3) * dact(z2)
    g1 = (g2 @ W2) * dact(z1)
    grad = pack(g1.T @ x, g1.sum(0), g2.T @ h1, g2.sum(0), g3.T @ h2, g3.sum(0), g4.T @ h3, g4.sum(0))
    return loss, grad


# ---------------------------------------------------------------------------
# BF16 rounding points of the device
# ---------------------------------------------------------------------------

def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split(a):
    """BF16 high and low parts of FP32 values."""
    a32 = np.asarray(a, dtype=np.float32)
    hi = bf16(a32)
    lo = bf16((a32.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


def forward_device(p, x):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    x = np.asarray(x, dtype=np.float32)
    xh, xl = split(x)
    W1h, W1l = split(W1)
    z1 = xh @ W1h.T + xl @ W1h.T + xh @ W1l.T + b1
    h1 = act(z1.astype(np.float32).astype(np.float64))
    z2 = bf16(h1) @ bf16(W2).T + b2
    h2 = act(z2.astype(np.float32).astype(np.float64))
    z3 = bf16(h2) @ bf16(W3).T + b3
    h3 = act(z3.astype(np.float32).astype(np.float64))
    h3h, h3l = split(h3)
    W4h, W4l = split(W4)
    o = h3h @ W4h.T + h3l @ W4h.T + h3h @ W4l.T + b4
    return o, (xh, z1, h1, z2, h2, z3, h3)


def grad_device(p, x, y):
    """The gradient the device training step computes: backward operands
    rounded to BF16 (gradients, transposed weights, stored activations)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    o, (xh, z1, h1, z2, h2, z3, h3) = forward_device(p, x)
    B = xh.shape[0]
    d = o - np.asarray(y, dtype=np.float64)
    loss = float((d * d).sum() / (6.0 * B))
    g4 = bf16(2.0 * d / (6.0 * B))
    g3 = bf16((g4 @ bf16(W4)) * dact(z3))
    g2 = bf16((g3 @ bf16(W3)) * dact(z2))
    g1 = bf16((g2 @ bf16(W2)) * dact(z1))
    grad = pack(g1.T @ xh, g1.sum(0), g2.T @ bf16(h1), g2.sum(0), g3.T @ bf16(h2), g3.sum(0),
                g4.T @ bf16(h3), g4.sum(0))
    return loss, grad


# ---------------------------------------------------------------------------
# Inputs, targets and the correction (FP64)
# ---------------------------------------------------------------------------

MU_EARTH = 398600.4418
R_EARTH = 6378.137


def raw_features(states, moon, tau):
    """The ten frame inputs and the frame of each state (n x 6), FP64."""
    s = np.asarray(states, dtype=np.float64)
    r, v = s[:, :3], s[:, 3:]
    rn = np.linalg.norm(r, axis=1)
    R = r / rn[:, None]
    h = np.cross(r, v)
    W = h / np.linalg.norm(h, axis=1)[:, None]
    S = np.cross(W, R)
    vc = np.sqrt(MU_EARTH / rn)
    w = vc / rn * tau
    m = np.asarray(moon, dtype=np.float64)
    M = m / np.linalg.norm(m)
    f = np.stack([10.0 * w, (v * R).sum(1) / vc, (v * S).sum(1) / vc - 1.0, R[:, 2], S[:, 2], W[:, 2],
                  R @ M, S @ M, W @ M, (R_EARTH / rn) ** 2], axis=1)
    return f, (R, S, W, rn * w ** 5, vc * w ** 5)


def correction(out, frame, norm):
    """Position and velocity corrections (n x 6) from standardised outputs."""
    R, S, W, sr, sv = frame
    c = np.asarray(out, dtype=np.float64) * np.asarray(norm["osig"], dtype=np.float64) + \
        np.asarray(norm["omu"], dtype=np.float64)
    dr = (c[:, 0:1] * R + c[:, 1:2] * S + c[:, 2:3] * W) * sr[:, None]
    dv = (c[:, 3:4] * R + c[:, 4:5] * S + c[:, 5:6] * W) * sv[:, None]
    return np.concatenate([dr, dv], axis=1)
