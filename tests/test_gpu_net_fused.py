"""The one-kernel network step of surrogate runs (features, the four layers
with on-chip activations and the increment in one strip march) against the
separate feature, convolution and increment kernels, which the tests of
test_gpu_net.py pin to the FP64 network restatement (oracle/net_oracle.py).

Both apply the increments of the same network to the same one-corrector
state.  They differ only where an activation's FP32 accumulation, summed in
another order, rounds to a neighbouring BF16 value, so the increments agree
to BF16 tolerance; everything outside the network (the predictor and
corrector, features, the increment's arithmetic) is the same code.
"""
import numpy as np
import pytest

from mass_hd_amd import lyo
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu


def _model(LX, LY, seed, npc=2):
    m = lyo.Lyotropic(LX, LY, npc=npc, precision="fp32", **CONFIG0_PHYS)
    m.configure(**dict(CONFIG0_INIT, seed=seed))
    return m


def _velocity(m):
    from oracle import diag_oracle as dg
    vx, vy = dg.velocity(m.get("ff"), m.LX, m.LY)
    return np.stack([vx, vy])


@pytest.fixture(scope="module")
def trained_net(require_gpu):
    net = lyo.Net(seed=17)
    m = _model(64, 64, 41)
    for _ in range(4):
        m.step(20)
        net.collect(m)
    m.close()
    net.train(200, lr=2e-3)
    return net


def _increments(LX, LY, net, fused, seed=7, before=30):
    """(one-corrector state, surrogate state) after one surrogate step from
    the state reached by `before` plain steps."""
    base = _model(LX, LY, seed)
    base.step(before)
    one = lyo.Lyotropic(LX, LY, npc=1, precision="fp32", **CONFIG0_PHYS)
    one.set_state(base.get("ff"), base.get("QQxx"), base.get("QQyx"), base.get("phi"))
    one.step(1)
    sur = _model(LX, LY, seed)
    sur.step(before - 1)
    lyo.enable_surrogate(sur, net, judge_every=1000000, tol=1.0)
    lyo.set_surrogate_kernels(sur, fused)
    sur.step(2)  # a judged plain step, then one surrogate step
    s = lyo.surrogate_stats(sur)
    assert s["surrogate_steps"] == 1 and s["judged"] == 1
    return one, sur


@pytest.mark.parametrize("LX,LY", [(64, 64), (16, 16), (37, 112), (130, 48), (20, 208)])
def test_fused_step_equals_separate_kernels(trained_net, LX, LY):
    """Strip edges (LY not a multiple of the 52 interior columns, a strip
    wider than the grid), chunk edges and the smallest grid: the fused step's
    increments of u, Q and phi equal the separate kernels' to BF16 tolerance,
    node by node."""
    one, sep = _increments(LX, LY, trained_net, 0)
    _, fus = _increments(LX, LY, trained_net, 1)
    for name in ("QQxx", "QQyx", "phi"):
        d_sep = sep.get(name) - one.get(name)
        d_fus = fus.get(name) - one.get(name)
        scale = np.abs(d_sep).max()
        assert scale > 0, name
        err = np.abs(d_fus - d_sep).max() / scale
        rms = np.sqrt(np.mean((d_fus - d_sep) ** 2)) / np.sqrt(np.mean(d_sep ** 2))
        print("%dx%d %s: max %.3g rms %.3g" % (LX, LY, name, err, rms))
        assert err < 3e-2, (name, err)
        assert rms < 1e-2, (name, rms)
    u1 = _velocity(one)
    du_sep = _velocity(sep) - u1
    du_fus = _velocity(fus) - u1
    err = np.abs(du_fus - du_sep).max() / np.abs(du_sep).max()
    print("%dx%d u: max %.3g" % (LX, LY, err))
    assert err < 3e-2, err
    # mass: the increment changes no node's density
    n_one = one.get("ff").sum(axis=-1)
    n_fus = fus.get("ff").sum(axis=-1)
    assert np.abs(n_fus - n_one).max() < 1e-4 * np.abs(n_one).max()


def test_fused_run_matches_separate_run(trained_net):
    """Thirty surrogate-mode steps with a judgement every ten: the same
    judgements, no fallbacks either way, judged errors alike, and states that
    stay within the BF16 tolerance of the network's increments."""
    runs = []
    for fused in (0, 1):
        m = _model(96, 96, 3)
        m.step(20)
        lyo.enable_surrogate(m, trained_net, judge_every=10, tol=1.0)
        lyo.set_surrogate_kernels(m, fused)
        m.step(30)
        runs.append((m, lyo.surrogate_stats(m)))
    (a, sa), (b, sb) = runs
    for k in ("steps", "surrogate_steps", "plain_steps", "judged", "fallbacks"):
        assert sa[k] == sb[k], k
    assert sa["surrogate_steps"] == 27 and sa["judged"] == 3
    assert sb["max_error"] == pytest.approx(sa["max_error"], rel=0.1)
    for name in ("QQxx", "phi"):
        x, y = a.get(name), b.get(name)
        assert np.abs(x - y).max() < 1e-4 * max(1.0, np.abs(x).max()), name
    fa, fb = a.sums(), b.sums()
    assert fb[0] == pytest.approx(fa[0], rel=1e-9)  # sum f: the increments keep every node's mass


def test_fused_kernels_switch(trained_net):
    m = _model(32, 32, 1)
    lyo.enable_surrogate(m, trained_net)
    lyo.set_surrogate_kernels(m, 0)
    lyo.set_surrogate_kernels(m, 1)
    with pytest.raises(lyo.MassLyoError):
        lyo.set_surrogate_kernels(m, 2)
    d = lyo.Lyotropic(32, 32, npc=2, precision="fp64", **CONFIG0_PHYS)
    with pytest.raises(lyo.MassLyoError):
        lyo.set_surrogate_kernels(d, 1)  # FP64 runs use the separate kernels
    small = lyo.Lyotropic(8, 16, npc=2, precision="fp32", **CONFIG0_PHYS)
    with pytest.raises(lyo.MassLyoError):
        lyo.set_surrogate_kernels(small, 1)
    lyo.set_surrogate_kernels(small, 0)


def _layers_scaled(net, factor):
    """the network's parameters with the last layer (weights and bias)
    multiplied by factor (a power of two: exact in BF16 and FP32)"""
    p = net.get_params().copy()
    n4 = 9 * 32 * 8 + 8
    p[-n4:] *= factor
    return p


@pytest.mark.parametrize("fused", [1, 0])
def test_calibration_is_scale_equivariant(trained_net, fused):
    """Output calibration fits a_c = sum t.p / sum p.p per output over the
    judged steps.  A network whose last layer is doubled predicts exactly
    twice the increments, so its factors come out exactly halved and, after
    the first judgement, the calibrated run applies bit for bit the original
    network's calibrated increments (same state, judgements and fallbacks);
    without calibration the two runs differ."""
    p0 = trained_net.get_params().copy()
    runs = {}
    try:
        for factor in (1.0, 2.0):
            trained_net.set_params(_layers_scaled(trained_net, factor) if factor != 1.0 else p0)
            for calib in (1, 0):
                m = _model(48, 64, 9)
                m.step(10)
                lyo.enable_surrogate(m, trained_net, judge_every=4, tol=1.0)
                lyo.set_surrogate_kernels(m, fused)
                lyo.set_surrogate_calibration(m, calib)
                m.step(13)
                runs[(factor, calib)] = (m.get("ff"), m.get("QQxx"), m.get("phi"),
                                         lyo.surrogate_stats(m), lyo.surrogate_calibration(m))
                m.close()
    finally:
        trained_net.set_params(p0)
    a1, a2 = runs[(1.0, 1)][4], runs[(2.0, 1)][4]
    assert all(0.25 < a < 4.0 for a in a1), a1
    assert any(a != 1.0 for a in a1), a1
    assert a2 == [a / 2 for a in a1], (a1, a2)
    for k in range(3):
        assert np.array_equal(runs[(1.0, 1)][k], runs[(2.0, 1)][k]), k
    # the first judgement measures the uncalibrated network (factors one);
    # from then on the two runs apply the same increments
    s1, s2 = runs[(1.0, 1)][3], runs[(2.0, 1)][3]
    for k in ("steps", "surrogate_steps", "plain_steps", "judged", "fallbacks", "last_error"):
        assert s1[k] == s2[k], k
    assert runs[(1.0, 0)][4] == [1.0] * 5
    assert not np.array_equal(runs[(1.0, 0)][1], runs[(2.0, 0)][1])
    assert not np.array_equal(runs[(1.0, 0)][1], runs[(1.0, 1)][1])


def test_calibration_switch_needs_surrogate(require_gpu):
    m = _model(32, 32, 1)
    with pytest.raises(lyo.MassLyoError):
        lyo.set_surrogate_calibration(m, 1)


@pytest.mark.parametrize("fused", [1, 0])
def test_surrogate_conserves_phi(trained_net, fused):
    """The plain corrector's phi increment is a flux divergence and keeps
    sum(phi); the network's is not, so surrogate steps remove its mean.  Over
    thirty steps in surrogate mode sum(phi) stays where the plain steps keep
    it (FP64 sums of the compensated FP32 phi), as does sum(f)."""
    m = _model(80, 96, 13)
    f0, p0 = m.sums()
    lyo.enable_surrogate(m, trained_net, judge_every=10, tol=1.0)
    lyo.set_surrogate_kernels(m, fused)
    m.step(30)
    assert lyo.surrogate_stats(m)["surrogate_steps"] == 27
    f1, p1 = m.sums()
    assert abs(p1 / p0 - 1.0) < 1e-11, p1 / p0 - 1.0
    assert abs(f1 / f0 - 1.0) < 1e-11, f1 / f0 - 1.0
