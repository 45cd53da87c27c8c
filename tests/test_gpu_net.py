"""Surrogate corrector network on the GPU: forward and gradients against the
FP64 restatement (oracle/net_oracle.py) to BF16 tolerance, training reduces
the loss, the trained network beats the method without it, and the judged
plain step takes every step back to the plain corrector when the tolerance
cannot be met.
"""
import numpy as np
import pytest

from mass_hd_amd import lyo
from oracle import net_oracle as no
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu


def oracle_params(net):
    return [[W.astype(np.float64), b.astype(np.float64)] for (W, b) in net.layers()]


def test_forward_matches_oracle(require_gpu):
    rng = np.random.default_rng(0)
    net = lyo.Net(seed=3)
    p = net.get_params()
    p += rng.normal(0, 0.05, p.shape).astype(np.float32)  # non-zero biases too
    net.set_params(p)
    x = rng.normal(size=(7, 32, 8)).astype(np.float32)
    got = net.forward(x)
    ref, _ = no.forward(x.astype(np.float64), oracle_params(net))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 3e-2, err  # BF16 weights and activations, FP32 accumulation


def test_inference_path_matches_forward(require_gpu):
    """Surrogate inference (FP32 features normalised inside the first layer's
    loads, output scaled inside the last layer) equals the plain forward pass
    bit for bit under the identity normalisation of a fresh network, on
    ragged row counts (tasks of 8 rows with a partial last one)."""
    rng = np.random.default_rng(2)
    net = lyo.Net(seed=4)
    p = net.get_params()
    p += rng.normal(0, 0.05, p.shape).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.normalization()[:8], np.zeros(8, np.float32))
    for LX, LY in ((7, 32), (9, 16), (16, 48)):
        x = rng.normal(size=(LX, LY, 8)).astype(np.float32)
        assert np.array_equal(net.predict(x), net.forward(x)), (LX, LY)


def test_gradients_match_oracle(require_gpu):
    rng = np.random.default_rng(1)
    net = lyo.Net(seed=5)
    p = net.get_params()
    p += rng.normal(0, 0.05, p.shape).astype(np.float32)
    net.set_params(p)
    x = rng.normal(size=(6, 16, 8)).astype(np.float32)
    y = rng.normal(size=(6, 16, 8)).astype(np.float32)
    L, g = net.grad(x, y)
    Lr, gr = no.loss_and_grads(x.astype(np.float64), y.astype(np.float64), oracle_params(net))
    assert L == pytest.approx(Lr, rel=3e-2)
    flat = np.concatenate([np.concatenate([a.reshape(-1), b]) for a, b in gr])
    cos = float(g @ flat / (np.linalg.norm(g) * np.linalg.norm(flat)))
    assert cos > 0.998, cos
    assert np.linalg.norm(g - flat) / np.linalg.norm(flat) < 6e-2
    o = 0
    for (W, b) in gr:  # per layer, too
        n = W.size + b.size
        ref = np.concatenate([W.reshape(-1), b])
        assert np.linalg.norm(g[o:o + n] - ref) / np.linalg.norm(ref) < 8e-2
        o += n


def _model(L, npc, seed):
    m = lyo.Lyotropic(L, L, npc=npc, precision="fp32", **CONFIG0_PHYS)
    m.configure(**dict(CONFIG0_INIT, seed=seed))
    return m


def _u(m):
    """velocity of the populations (not the frame's derived ux/uy, which are
    the inputs of the last corrector)"""
    from oracle import diag_oracle as dg
    vx, vy = dg.velocity(m.get("ff"), m.LX, m.LY)
    return np.concatenate([vx.reshape(-1), vy.reshape(-1)])


def test_training_reduces_loss_and_beats_no_network(require_gpu):
    L = 64
    net = lyo.Net(seed=11)
    for seed in (21, 22):  # training frames from plain npc=2 runs
        m = _model(L, 2, seed)
        for _ in range(4):
            m.step(25)
            net.collect(m)
        m.close()
    assert net.dataset_size == 8
    l0, l1 = net.train(300, lr=2e-3)
    assert l1 < 0.5 * l0, (l0, l1)
    # inference with the trained statistics against the oracle
    nrm = net.normalization().astype(np.float64)
    xr = np.random.default_rng(3).normal(size=(5, 16, 8))
    x = (xr / nrm[8:16] + nrm[:8]).astype(np.float32)
    got = net.predict(x)
    ref, _ = no.forward((x.astype(np.float64) - nrm[:8]) * nrm[8:16], oracle_params(net))
    ref = ref * nrm[16:24]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 3e-2, err

    # From one state S (a run after 40 steps), one more step three ways:
    # (a) plain npc=2 -- the reference; (b) predictor + one corrector with no
    # network; (c) predictor + one corrector + the network.  Surrogate mode
    # judges its first step, so (c) enables it one step earlier.
    ref = _model(L, 2, 31)
    ref.step(41)
    u_ref = _u(ref)
    base = _model(L, 2, 31)
    base.step(40)
    one = lyo.Lyotropic(L, L, npc=1, precision="fp32", **CONFIG0_PHYS)
    one.set_state(base.get("ff"), base.get("QQxx"), base.get("QQyx"), base.get("phi"))
    one.step(1)
    sur = _model(L, 2, 31)
    sur.step(39)
    lyo.enable_surrogate(sur, net, judge_every=1000000, tol=1.0)
    sur.step(2)  # a judged plain step, then one surrogate step
    s = lyo.surrogate_stats(sur)
    assert s["surrogate_steps"] == 1 and s["judged"] == 1
    e_one = np.linalg.norm(_u(one) - u_ref) / np.linalg.norm(u_ref)
    e_sur = np.linalg.norm(_u(sur) - u_ref) / np.linalg.norm(u_ref)
    assert e_sur < 0.7 * e_one, (e_sur, e_one)


def test_judge_falls_back_to_plain_steps(require_gpu):
    """With a tolerance no prediction can meet, every judgement fails and
    every step is plain: the run equals a plain npc=2 run bit for bit."""
    L = 32
    net = lyo.Net(seed=2)
    m = _model(L, 2, 5)
    m.step(3)
    net.collect(m)
    net.train(5)
    a = _model(L, 2, 8)
    lyo.enable_surrogate(a, net, judge_every=4, tol=1e-300)
    a.step(13)
    s = lyo.surrogate_stats(a)
    assert s["surrogate_steps"] == 0 and s["plain_steps"] == 13
    assert s["judged"] == 4 and s["fallbacks"] == 4
    b = _model(L, 2, 8)
    b.step(13)
    for k in ("ff", "QQxx", "phi"):
        assert np.array_equal(a.get(k), b.get(k))


def test_surrogate_rejects_bad_setups(require_gpu):
    net = lyo.Net(seed=1)
    m = _model(32, 2, 1)
    with pytest.raises(lyo.MassLyoError):
        lyo.enable_surrogate(m, net)  # untrained: no normalisation
    one = _model(32, 1, 1)
    with pytest.raises(lyo.MassLyoError):
        net.collect(one)  # npc=1 has no one-corrector/two-corrector pair
