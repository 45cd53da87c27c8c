"""Walled grids (bc 1-4) on the GPU against the CPU oracle.

The ghost-layer kernels (mass_hd_amd/csrc/lyo_walls.hpp) follow the
reference's loop structure in FP64, so every domain field must equal the
oracle's bit for bit: initial fields (including the wall-specific initial
configurations), every frame field after one step for npc 1-3 and grids down
to 2x2, and the state after 60 steps.  At 512^2 the size-independent checks
hold: mass and phi conserved, the run finite and deterministic.
"""
import math

import numpy as np
import pytest

from mass_hd_amd import lyo
from oracle import pyoracle as po
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

pytestmark = pytest.mark.gpu

DERIVED = ["n", "ux", "uy", "sigmaXX", "sigmaXY", "sigmaYX", "sigmaYY", "sigma_bulk",
           "sigma_elastic_xx", "sigma_elastic_yx", "sigma_phase_field_xx",
           "sigma_phase_field_yx", "sigma_active_xx", "sigma_active_yx", "MU", "HHxx", "HHyx"]
STATE = ["ff", "QQxx", "QQyx", "phi"]
ACTIVE = dict(CONFIG0_PHYS, zeta=0.05)
BCS = [1, 2, 3, 4]


def pair(LX, LY, bc, npc=1, phys=ACTIVE, init=CONFIG0_INIT, loops=False):
    g = lyo.Lyotropic(LX, LY, npc=npc, precision=lyo.FP64, bc=bc, **phys)
    if loops:
        g.set_loop_kernels(True)
    o = po.Oracle(LX, LY, npc=npc, bc=bc, **phys)
    fg = g.configure(**init)
    fo = o.configure(**init)
    return g, o, fg, fo


def assert_bitwise(g, o, names, what=""):
    for n in names:
        a, b = g.get(n), o.get(n)
        assert np.array_equal(a, b), "%s %s differs: max |d| = %g" % (what, n, np.abs(a - b).max())


@pytest.mark.parametrize("loops", [False, True])
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("npc", [1, 2, 3])
@pytest.mark.parametrize("LX,LY", [(2, 2), (5, 3), (16, 12), (45, 67), (17, 33), (23, 130)])
def test_walls_one_step_bit_exact(require_gpu, bc, npc, LX, LY, loops):
    # loops: the node loops with condition kernels between them; otherwise
    # the fused sweeps (10 x 62 tiles: 17 x 33 leaves partial tiles, 45 x 67
    # spans two tiles in y, 23 x 130 three, with an interior y-tile)
    print("walls one step bc=%d npc=%d %dx%d loops=%d" % (bc, npc, LX, LY, loops), flush=True)
    g, o, fg, fo = pair(LX, LY, bc, npc, loops=loops)
    assert fg == fo
    assert_bitwise(g, o, STATE, "initial")
    g.step(1)
    o.step(1)
    assert_bitwise(g, o, STATE, "step 1")
    assert_bitwise(g, o, DERIVED, "step 1")


@pytest.mark.parametrize("loops", [False, True])
@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("LX,LY", [(70, 52), (23, 130)])
def test_walls_60_steps_bit_exact(require_gpu, bc, loops, LX, LY):
    print("walls 60 steps bc=%d loops=%d %dx%d" % (bc, loops, LX, LY), flush=True)
    g, o, _, _ = pair(LX, LY, bc, npc=2, loops=loops)
    g.step(60)
    o.step(60)
    assert_bitwise(g, o, STATE + ["ux", "sigmaXY", "MU"])
    ux = g.get("ux")
    assert np.abs(ux).max() > 1e-5  # an active flow, not a rest state


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("LX,LY", [(33, 40), (23, 130)])
def test_walls_switching_kernel_sets_mid_run(require_gpu, bc, LX, LY):
    # fused sweeps and node loops hand the state over at step boundaries
    g, o, _, _ = pair(LX, LY, bc, npc=2)
    for s, loops in enumerate([False, True, True, False, True, False]):
        g.set_loop_kernels(loops)
        g.step(s % 3 + 1)
        o.step(s % 3 + 1)
    assert_bitwise(g, o, STATE + ["ux", "sigmaXY", "MU"])


@pytest.mark.parametrize("config", ["boxatwall", "cuttingatwall", "circleatwall", "wettedwall"])
@pytest.mark.parametrize("bc", [2, 4])
def test_wall_initial_configurations(require_gpu, config, bc):
    # bc 4 takes the wall branches of boxatwall and cuttingatwall
    # (lyotropic.cpp:96-116), every other case the periodic ones
    init = dict(config=config, seed=3, angle=0.0, noise=0.5, qamp=1.0, level=8.0, radius=6.0)
    g, o, fg, fo = pair(40, 36, bc, npc=1, phys=dict(ACTIVE, conc=1.0), init=init)
    assert fg == fo
    assert_bitwise(g, o, STATE)
    assert g.get("phi").any()
    g.step(3)
    o.step(3)
    assert_bitwise(g, o, STATE)


@pytest.mark.parametrize("bc", BCS)
def test_walls_sums_and_conservation(require_gpu, bc):
    g, o, _, _ = pair(64, 48, bc, npc=2)
    f0, p0 = g.sums()
    g.step(40)
    o.step(40)
    fs, ps = g.sums()
    assert fs == pytest.approx(math.fsum(o.get("ff").reshape(-1)), rel=4e-15)
    assert ps == pytest.approx(math.fsum(o.get("phi")), rel=4e-15)
    assert fs == pytest.approx(f0, rel=1e-13) and ps == pytest.approx(p0, rel=1e-12)
    g.runtime_checks()


@pytest.mark.parametrize("bc", [2, 4])
def test_walls_512_conserves_and_repeats(require_gpu, bc):
    print("walls 512^2 bc=%d" % bc, flush=True)
    runs = []
    for _ in range(2):
        g = lyo.Lyotropic(512, 512, npc=2, precision=lyo.FP64, bc=bc, **ACTIVE)
        g.configure(**CONFIG0_INIT)
        f0, p0 = g.sums()
        g.step(50)
        f1, p1 = g.sums()
        assert f1 == pytest.approx(f0, rel=1e-13) and p1 == pytest.approx(p0, rel=1e-12)
        runs.append([g.get(n) for n in STATE])
        assert all(np.isfinite(a).all() for a in runs[-1])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("bc", [1, 4])
def test_walls_checkpoint_continues_bitwise(require_gpu, bc):
    LX, LY = 40, 30
    a = lyo.Lyotropic(LX, LY, npc=2, precision=lyo.FP64, bc=bc, **ACTIVE)
    a.configure(**CONFIG0_INIT)
    a.step(4)
    img = a.checkpoint()
    assert len(img) == 64 + 12 * (LX + 2) * (LY + 2) * 8
    a.step(5)
    b = lyo.Lyotropic(LX, LY, npc=2, precision=lyo.FP64, bc=bc, **ACTIVE)
    b.restore(img)
    b.step(5)
    for n in STATE + ["ux", "sigmaXY"]:
        assert np.array_equal(a.get(n), b.get(n)), n
    # another boundary condition, or a periodic run, refuses the image
    for other in (0, 3 if bc != 3 else 2):
        c = lyo.Lyotropic(LX, LY, npc=2, precision=lyo.FP64, bc=other, **ACTIVE)
        with pytest.raises(lyo.MassLyoError):
            c.restore(img)


@pytest.mark.parametrize("bc", [1, 2, 3, 4])
def test_walls_fused_sweeps_equal_loops_large(require_gpu, bc):
    """The fused tile sweeps against the reference's node loops on a large
    grid of odd sizes (1003 x 517: partial tiles in both directions), npc 2,
    every state and frame field bit for bit."""
    print("walls bc=%d fused vs loops 1003x517" % bc, flush=True)
    runs = []
    for loops in (False, True):
        m = lyo.Lyotropic(1003, 517, npc=2, precision=lyo.FP64, bc=bc,
                          **dict(CONFIG0_PHYS, zeta=0.03))
        m.set_loop_kernels(loops)
        m.configure(**CONFIG0_INIT)
        m.step(4)
        runs.append({f: m.get(f) for f in STATE + DERIVED})
        m.close()
    for f, v in runs[0].items():
        assert np.array_equal(v, runs[1][f]), f


def test_walls_refuse_what_needs_a_periodic_grid(require_gpu):
    with pytest.raises(lyo.MassLyoError):
        lyo.Lyotropic(16, 16, npc=2, precision=lyo.FP32, bc=2, **ACTIVE)
    with pytest.raises(lyo.MassLyoError):
        lyo.Lyotropic(16, 16, npc=2, precision=lyo.FP64, bc=5, **ACTIVE)
    g = lyo.Lyotropic(16, 16, npc=2, precision=lyo.FP64, bc=3, **ACTIVE)
    g.configure(**CONFIG0_INIT)
    with pytest.raises(lyo.MassLyoError):
        g.diagnostics()
    g.set_loop_kernels(True)
    g.set_loop_kernels(False)
    run = dict(LX=16, LY=16, bc=2, **ACTIVE)
    with pytest.raises(lyo.MassLyoError):
        lyo.Batch([run], npc=2, precision=lyo.FP64)
