"""Surrogate corrector network on the GPU: forward and gradients against the
FP64 restatement (oracle/net_oracle.py) to BF16 tolerance, training reduces
the loss, the trained network beats the method without it, and the judge
hands every step back to the plain corrector when the tolerance cannot be met.
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
    # per layer, too
    o = 0
    for (W, b) in gr:
        n = W.size + b.size
        seg = g[o:o + n]
        ref = np.concatenate([W.reshape(-1), b])
        assert np.linalg.norm(seg - ref) / np.linalg.norm(ref) < 8e-2
        o += n


def _model(L, npc, seed, precision="fp32"):
    m = lyo.Lyotropic(L, L, npc=npc, precision=precision, **CONFIG0_PHYS)
    m.configure(**dict(CONFIG0_INIT, seed=seed))
    return m


def test_training_reduces_loss_and_beats_no_network(require_gpu):
    L = 64
    net = lyo.Net(seed=11)
    # training frames from plain npc=2 runs
    for seed in (21, 22):
        m = _model(L, 2, seed)
        for _ in range(4):
            m.step(25)
            net.collect(m)
        m.close()
    assert net.dataset_size == 8
    l0, l1 = net.train(300, lr=2e-3)
    assert l1 < 0.5 * l0, (l0, l1)

    # gain: from one state, one step of (a) npc=2 (reference), (b) npc=1,
    # (c) npc=1 + network
    def one_step(mode):
        m = _model(L, 2, 31)
        m.step(40)
        if mode == "plain":
            m.step(1)
        elif mode == "surrogate":
            lyo.enable_surrogate(m, net, judge_every=1000000, tol=1.0)
            m.step(1)  # step 0 of surrogate mode is judged (plain): skip it
        u = m.get("ux"), m.get("uy")
        m.close()
        return np.concatenate(u)
    # surrogate mode's first step is a judged plain step, so compare on the
    # step after: build the states explicitly
    m_ref = _model(L, 2, 31)
    m_ref.step(41)
    u_ref = np.concatenate([m_ref.get("ux"), m_ref.get("uy")])
    m_sur = _model(L, 2, 31)
    m_sur.step(40)
    lyo.enable_surrogate(m_sur, net, judge_every=1000000, tol=1.0)
    m_sur.step(1)   # judged plain step (count 0)
    m_sur2 = _model(L, 2, 31)
    m_sur2.step(39)
    lyo.enable_surrogate(m_sur2, net, judge_every=1000000, tol=1.0)
    m_sur2.step(2)  # judged plain step, then one surrogate step
    s = lyo.surrogate_stats(m_sur2)
    assert s["surrogate_steps"] == 1 and s["judged"] == 1
    u_sur = np.concatenate([m_sur2.get("ux"), m_sur2.get("uy")])
    m_one = _model(L, 2, 31)
    m_one.step(40)
    # npc=1 step on the same state: a fresh npc=1 run stepped to the same point
    m_np1 = lyo.Lyotropic(L, L, npc=1, precision="fp32", **CONFIG0_PHYS)
    m_np1.set_state(m_one.get("ff"), m_one.get("QQxx"), m_one.get("QQyx"), m_one.get("phi"))
    m_np1.step(1)
    u_np1 = np.concatenate([m_np1.get("ux"), m_np1.get("uy")])
    e_np1 = np.linalg.norm(u_np1 - u_ref) / np.linalg.norm(u_ref)
    e_sur = np.linalg.norm(u_sur - u_ref) / np.linalg.norm(u_ref)
    assert e_sur < 0.7 * e_np1, (e_sur, e_np1)


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
