"""Batched runs: each run of a batch equals the same run on its own handle,
bit for bit, whatever the launch shape (persistent grid, chunk rows, graph
replays and the odd-step tail) and whatever the step's launch plan (one-kernel
step for FP32 npc=2, predictor + fused pair, separate sweeps)."""
import numpy as np
import pytest

from mass_hd_amd import lyo
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu

RUNS = [
    dict(CONFIG0_PHYS, LX=64, LY=96, zeta=0.02, xi=0.7),
    dict(CONFIG0_PHYS, LX=128, LY=70, zeta=0.05, xi=1.0),
    dict(CONFIG0_PHYS, LX=37, LY=53, zeta=0.001, xi=0.5),
]


def solo(run, npc, precision, seed, steps):
    r = dict(run)
    LX, LY = r.pop("LX"), r.pop("LY")
    m = lyo.Lyotropic(LX, LY, npc=npc, precision=precision, **r)
    m.configure(**dict(CONFIG0_INIT, seed=seed))
    m.step(steps)
    out = {k: m.get(k) for k in ("ff", "QQxx", "QQyx", "phi", "sigmaXY")}
    m.close()
    return out


@pytest.mark.parametrize("precision,npc", [("fp64", 2), ("fp32", 2), ("fp32", 1), ("fp32", 3),
                                           ("fp64", 3)])
@pytest.mark.parametrize("launch", [(0, 0), (4, 64), (16, 0)])
def test_batch_equals_solo_runs(require_gpu, precision, npc, launch):
    steps = 7
    b = lyo.Batch(RUNS, npc=npc, precision=precision)
    if launch != (0, 0):
        b.set_launch(*launch)
    for i in range(len(RUNS)):
        b.configure(i, **dict(CONFIG0_INIT, seed=100 + i))
    b.step(3)
    b.step(steps - 3)  # odd start parity: single step + graph replays
    for i, run in enumerate(RUNS):
        ref = solo(run, npc, precision, 100 + i, steps)
        for k, v in ref.items():
            assert np.array_equal(b.get(i, k), v), (i, k)
        assert b.sums(i)[0] == pytest.approx(v.size and ref["ff"].sum(), rel=1e-12)
    b.close()


def test_batch_autotune_then_run(require_gpu):
    b = lyo.Batch(RUNS, npc=2, precision="fp32")
    chunk, grid = b.autotune()
    assert chunk in (16, 24, 32, 48, 64, 96, 128)  # candidates of the one-kernel step
    for i in range(len(RUNS)):
        b.configure(i, **dict(CONFIG0_INIT, seed=7 + i))
    b.step(4)
    for i, run in enumerate(RUNS):
        ref = solo(run, 2, "fp32", 7 + i, 4)
        assert np.array_equal(b.get(i, "ff"), ref["ff"])
    with pytest.raises(lyo.MassLyoError):
        b.autotune()  # not allowed once runs hold state
    d = b.diagnostics(1)
    assert d["u_rms"] > 0
