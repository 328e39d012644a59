"""GPU tier: the learned corrector's matrix-core kernels against the oracle
(oracle/net_oracle.py).

Tolerances: the device accumulates in FP32 (the oracle in FP64) and uses the
hardware exp2 / reciprocal in SiLU; either can move a value across a BF16
rounding boundary (2^-9 relative). Outputs are standardised (O(1)).
"""
import numpy as np
import pytest

from oracle import net_oracle as no

pytestmark = pytest.mark.gpu


def test_inference_matches_oracle(engine):
    p = no.init_params(11)
    engine.net_set(p, {"fmu": np.zeros(10), "fsig": np.ones(10), "omu": np.zeros(6), "osig": np.ones(6),
                       "tau": 60.0})
    rng = np.random.default_rng(12)
    x = rng.normal(size=(1000, 10)).astype(np.float32)  # not a multiple of 64: partial tiles
    got = engine.net_infer(x)
    ref, _ = no.forward_device(p, x)
    exact, _ = no.forward_exact(p, x)
    err = np.abs(got - ref)
    print(f"inference vs oracle: max {err.max():.3e} median {np.median(err):.3e}; "
          f"vs FP64 {np.abs(got - exact).max():.3e}")
    assert np.median(err) < 1e-4 and err.max() < 5e-3
    assert np.abs(got - exact).max() < 2e-2


def test_gradient_matches_oracle(engine):
    p = no.init_params(13)
    engine.net_set(p, {"fmu": np.zeros(10), "fsig": np.ones(10), "omu": np.zeros(6), "osig": np.ones(6),
                       "tau": 60.0})
    rng = np.random.default_rng(14)
    x = rng.normal(size=(2048, 10)).astype(np.float32)
    y = rng.normal(size=(2048, 6)).astype(np.float32)
    g, loss = engine.net_gradient(x, y)
    l_ref, g_ref = no.grad_device(p, x, y)
    l_ex, g_ex = no.loss_and_grad_exact(p, x, y)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    rel_ex = np.linalg.norm(g - g_ex) / np.linalg.norm(g_ex)
    print(f"loss {loss:.6f} (oracle {l_ref:.6f}); gradient rel. error vs oracle {rel:.3e}, vs FP64 {rel_ex:.3e}")
    assert abs(loss - l_ref) <= 1e-4 * l_ref
    assert rel < 5e-3
    assert rel_ex < 3e-2
    # Every tensor, not only the largest ones, must agree.
    for a, b in ((no.OFF_W1, no.OFF_B1), (no.OFF_B1, no.OFF_W2), (no.OFF_W4, no.OFF_B4), (no.OFF_B4, no.NPARAM)):
        r = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        assert r < 2e-2, (a, b, r)
