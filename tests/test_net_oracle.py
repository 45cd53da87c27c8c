"""The FP64 network restatement (oracle/net_oracle.py) against finite
differences and closed forms: its gradients are the checker of the GPU
backward kernels."""
import numpy as np
import pytest

from oracle import net_oracle as no


def test_gelu_closed_forms():
    assert no.gelu(0.0) == 0.0
    assert no.gelu(10.0) == pytest.approx(10.0)
    assert abs(no.gelu(-10.0)) < 1e-12
    x = np.linspace(-4, 4, 41)
    h = 1e-6
    assert np.allclose(no.gelu_grad(x), (no.gelu(x + h) - no.gelu(x - h)) / (2 * h), atol=1e-8)


def test_conv_is_a_periodic_correlation():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(5, 6, 3))
    W = rng.normal(size=(9, 2, 3))
    b = rng.normal(size=2)
    out = no.conv(a, W, b)
    x, y = 4, 0  # corner: wraps both ways
    ref = b.copy()
    for t, (dx, dy) in enumerate(no.TAPS):
        ref += W[t] @ a[(x + dx) % 5, (y + dy) % 6]
    assert np.allclose(out[x, y], ref)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(6, 5, 8))
    y = rng.normal(size=(6, 5, 8))
    ps = no.init_params(3)
    for l in range(4):
        ps[l][1] = rng.normal(0, 0.1, ps[l][1].shape)
    L, g = no.loss_and_grads(x, y, ps)
    h = 1e-6
    for l in range(4):
        for (idx, which) in [((rng.integers(9), rng.integers(ps[l][0].shape[1]),
                               rng.integers(ps[l][0].shape[2])), 0), ((rng.integers(ps[l][1].size),), 1)]:
            for _ in range(2):
                p = ps[l][which]
                old = p[idx]
                p[idx] = old + h
                Lp, _ = no.loss_and_grads(x, y, ps)
                p[idx] = old - h
                Lm, _ = no.loss_and_grads(x, y, ps)
                p[idx] = old
                fd = (Lp - Lm) / (2 * h)
                assert g[l][which][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), (l, which, idx)
