"""CPU tier: pins of the corrector's oracle (oracle/net_oracle.py).

The analytic backward pass of the FP64 restatement is checked against
central finite differences of the FP64 loss; the BF16 restatement (the
rounding points of the device) must stay close to the exact gradient; the
BF16 rounding itself is checked against known bit patterns.
"""
import numpy as np

from oracle import net_oracle as no


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, no.NIN)), rng.normal(size=(n, no.NOUT))


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39, 65504.0], dtype=np.float32)
    got = no.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0);
    # 1 + 3 * 2^-9 rounds up to 1 + 2^-7.
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7 and got[3] == -2.5
    assert got[5] == 65536.0
    hi, lo = no.split(np.float32(np.pi))
    assert abs(hi + lo - np.pi) < 2e-5 and abs(hi - np.pi) > 1e-4


def test_exact_gradient_matches_finite_differences():
    p = no.init_params(3).astype(np.float64)
    x, y = _batch(32, 4)
    _, g = no.loss_and_grad_exact(p, x, y)
    rng = np.random.default_rng(5)
    # A sample of coordinates from every tensor (weights and biases of each layer).
    bounds = [no.OFF_W1, no.OFF_B1, no.OFF_W2, no.OFF_B2, no.OFF_W3, no.OFF_B3, no.OFF_W4, no.OFF_B4, no.NPARAM]
    idx = np.concatenate([rng.integers(a, b, 6) for a, b in zip(bounds[:-1], bounds[1:])])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (no.loss_and_grad_exact(pp, x, y)[0] - no.loss_and_grad_exact(pm, x, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g).max()) + 1e-4 * abs(g[k]), (k, fd, g[k])


def test_device_rounding_gradient_close_to_exact():
    p = no.init_params(7)
    x, y = _batch(2048, 8)
    l_ex, g_ex = no.loss_and_grad_exact(p, x.astype(np.float32), y)
    l_bf, g_bf = no.grad_device(p, x, y)
    assert abs(l_bf - l_ex) <= 1e-3 * l_ex
    rel = np.linalg.norm(g_bf - g_ex) / np.linalg.norm(g_ex)
    assert rel < 2e-2, rel
    o_ex, _ = no.forward_exact(p, x.astype(np.float32))
    o_bf, _ = no.forward_device(p, x)
    assert np.abs(o_bf - o_ex).max() < 2e-2


def test_hi_lo_split_keeps_input_layer_accurate():
    # Inputs that BF16 alone would round by 2^-9 relative are carried to
    # about 2^-16 by the hi/lo split of the input layer.
    p = no.init_params(9)
    x = np.full((4, no.NIN), 1.0 + 2 ** -10)
    z_exact = x @ no.unpack(p)[0].T
    xh, xl = no.split(x)
    W1h, W1l = no.split(no.unpack(p)[0])
    z_split = xh @ W1h.T + xl @ W1h.T + xh @ W1l.T
    z_plain = no.bf16(x) @ no.bf16(no.unpack(p)[0]).T
    assert np.abs(z_split - z_exact).max() < np.abs(z_plain - z_exact).max() / 20
