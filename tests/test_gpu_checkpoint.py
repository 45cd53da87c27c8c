"""Binary checkpoints: a run continued from a checkpoint is bitwise the run
that was not interrupted (FP64 and FP32, with the FP32 phi compensation term),
images of another grid, precision, rho or format are refused with an error
code, and the host program restarts from a checkpoint file."""
import numpy as np
import pytest

from mass_hd_amd import lyo
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu

STATE = ["ff", "QQxx", "QQyx", "phi"]
DERIVED = ["ux", "sigmaXY", "MU", "sigma_active_xx"]


def _run(LX, LY, npc, precision, **phys):
    p = dict(CONFIG0_PHYS)
    p.update(phys)
    return lyo.Lyotropic(LX, LY, npc=npc, precision=precision, **p)


@pytest.mark.parametrize("precision,npc", [(lyo.FP64, 1), (lyo.FP64, 2), (lyo.FP32, 1),
                                           (lyo.FP32, 2), (lyo.FP32, 3)])
def test_checkpoint_continues_bitwise(require_gpu, precision, npc):
    LX, LY = 70, 64
    a = _run(LX, LY, npc, precision)
    a.configure(**CONFIG0_INIT)
    a.step(5)
    img = a.checkpoint()
    planes = 12 if precision == lyo.FP64 else 13
    size = 4 if precision == lyo.FP32 else 8
    assert img[:8] == b"MASSLYO1"
    assert len(img) == 64 + planes * LX * LY * size
    a.step(6)
    b = _run(LX, LY, npc, precision)
    b.restore(img)
    assert not np.any(b.get("sigmaXY"))  # derived fields restart at zero
    b.step(6)
    for n in STATE + DERIVED:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x, y), "%s differs: max |d| = %g" % (n, np.abs(x - y).max())
    assert a.sums() == b.sums()


def test_checkpoint_npc_may_differ(require_gpu):
    a = _run(64, 48, 2, lyo.FP32)
    a.configure(**CONFIG0_INIT)
    a.step(3)
    img = a.checkpoint()
    b = _run(64, 48, 1, lyo.FP32)
    b.restore(img)
    for n in STATE:
        assert np.array_equal(a.get(n), b.get(n))


def test_checkpoint_refuses_other_runs(require_gpu):
    a = _run(64, 48, 1, lyo.FP32)
    a.configure(**CONFIG0_INIT)
    img = a.checkpoint()
    cases = [(_run(48, 64, 1, lyo.FP32), img, "grid"),
             (_run(64, 48, 1, lyo.FP64), img, "precision"),
             (_run(64, 48, 1, lyo.FP32, rho=20.0), img, "rho"),
             (_run(64, 48, 1, lyo.FP32), img[:-4], "truncated"),
             (_run(64, 48, 1, lyo.FP32), b"XXXXXXXX" + img[8:], "not a checkpoint"),
             (_run(64, 48, 1, lyo.FP32), img[:40], "not a checkpoint")]
    for m, data, what in cases:
        with pytest.raises(lyo.MassLyoError) as e:
            m.restore(data)
        assert e.value.code == 1 and what in str(e.value), (what, str(e.value))
    fresh = _run(16, 16, 1, lyo.FP32)
    with pytest.raises(lyo.MassLyoError):  # nothing to save before configure/set_state
        fresh.checkpoint()


def test_fp64_restart_matches_the_oracle(require_gpu):
    """A restarted FP64 run against the CPU oracle's uninterrupted run: the
    checkpoint image holds the oracle's state bit for bit (SoA planes f v,
    Qxx, Qyx, phi after the header), and two restarts later the state and the
    frame fields still equal the oracle's."""
    from oracle import pyoracle as po
    LX, LY, npc = 52, 70, 2
    o = po.Oracle(LX, LY, npc=npc, **CONFIG0_PHYS)
    o.configure(**CONFIG0_INIT)
    a = _run(LX, LY, npc, lyo.FP64)
    a.configure(**CONFIG0_INIT)
    a.step(5)
    o.step(5)
    img = a.checkpoint()
    N = LX * LY
    planes = np.frombuffer(img, dtype=np.float64, offset=64, count=12 * N).reshape(12, N)
    assert np.array_equal(planes[:9].T, o.get("ff"))
    assert np.array_equal(planes[9], o.get("QQxx"))
    assert np.array_equal(planes[10], o.get("QQyx"))
    assert np.array_equal(planes[11], o.get("phi"))
    a.close()
    b = _run(LX, LY, npc, lyo.FP64)
    b.restore(img)
    b.step(4)
    img2 = b.checkpoint()
    b.close()
    c = _run(LX, LY, npc, lyo.FP64)
    c.restore(img2)
    c.step(3)
    o.step(7)
    for n in STATE + DERIVED:
        x, y = c.get(n), o.get(n)
        assert np.array_equal(x, y), "%s differs: max |d| = %g" % (n, np.abs(x - y).max())
