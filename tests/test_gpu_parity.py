"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

FP64 is the reference arithmetic in the reference order, so it must match the
oracle bit for bit (initial fields, every frame field after one step and after
100 steps).  FP32 is checked against the FP64 oracle with stated tolerances.
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


def pair(LX, LY, npc=1, precision=lyo.FP64, phys=CONFIG0_PHYS, init=CONFIG0_INIT):
    g = lyo.Lyotropic(LX, LY, npc=npc, precision=precision, **phys)
    o = po.Oracle(LX, LY, npc=npc, **phys)
    fg = g.configure(**init)
    fo = o.configure(**init)
    return g, o, fg, fo


def assert_bitwise(g, o, names):
    for n in names:
        a, b = g.get(n), o.get(n)
        assert np.array_equal(a, b), "%s differs: max |d| = %g" % (n, np.abs(a - b).max())


@pytest.mark.parametrize("LX,LY", [(128, 128), (37, 53), (2, 2), (3, 130), (64, 61)])
def test_initial_fields_bit_exact(require_gpu, LX, LY):
    g, o, fg, fo = pair(LX, LY)
    assert fg == fo  # ftot, ptot exactly as the reference computes them
    assert_bitwise(g, o, STATE)
    # derived fields are zero before the first step (Field value-initialisation)
    for n in ("sigmaXX", "ux", "MU"):
        assert not np.any(g.get(n))


@pytest.mark.parametrize("npc", [1, 2, 3])
@pytest.mark.parametrize("LX,LY", [(128, 128), (45, 67), (5, 3)])
def test_fp64_one_step_bit_exact(require_gpu, LX, LY, npc):
    g, o, _, _ = pair(LX, LY, npc=npc)
    g.step(1)
    o.step(1)
    assert_bitwise(g, o, STATE)
    # frame fields come from the last UpdateQuantities (start of the last corrector)
    assert_bitwise(g, o, DERIVED)


def test_fp64_config0_100_steps(require_gpu):
    """128^2, npc=1: bit-exact after 100 steps (the contract allows 1e-11)."""
    g, o, _, _ = pair(128, 128, npc=1)
    g.step(100)
    o.step(100)
    for n in STATE + ["sigmaXX", "sigmaXY", "ux"]:
        a, b = g.get(n), o.get(n)
        rel = np.abs(a - b).max() / np.abs(b).max()
        assert rel <= 1e-11, (n, rel)
        assert np.array_equal(a, b), (n, rel)


def test_fp64_npc2_50_steps_ragged(require_gpu):
    g, o, _, _ = pair(97, 131, npc=2)
    g.step(50)
    o.step(50)
    assert_bitwise(g, o, STATE)


@pytest.mark.parametrize("chunk,wpb", [(1, 1), (3, 2), (7, 4), (128, 4)])
def test_launch_shape_does_not_change_results(require_gpu, chunk, wpb):
    ref, _, _, _ = pair(70, 125, npc=2)
    ref.step(3)
    g, _, _, _ = pair(70, 125, npc=2)
    g.set_launch(chunk, wpb, chunk % 2)
    g.step(3)
    for n in STATE:
        assert np.array_equal(g.get(n), ref.get(n))
    g32, _, _, _ = pair(70, 125, npc=2, precision=lyo.FP32)
    r32, _, _, _ = pair(70, 125, npc=2, precision=lyo.FP32)
    g32.set_launch(chunk, wpb, chunk % 2)
    g32.step(3)
    r32.step(3)
    for n in STATE:
        assert np.array_equal(g32.get(n), r32.get(n))


@pytest.mark.parametrize("precision", [lyo.FP64, lyo.FP32])
@pytest.mark.parametrize("npc", [2, 3])
@pytest.mark.parametrize("LX,LY,xc2", [(128, 128, 32), (45, 67, 7), (5, 3, 1), (2, 60, 4),
                                       (70, 113, 64), (40, 109, 16), (24, 54, 5)])
def test_fused_pair_equals_separate_sweeps(require_gpu, precision, npc, LX, LY, xc2):
    """The last two correctors in one kernel give the separate sweeps' results
    bit for bit, for the state and for the frame's derived fields (which the
    fused path recomputes on demand); asking for derived fields mid-run does
    not disturb the run."""
    runs = []
    # (fused pair, march order, one-kernel step): separate sweeps, all chunks
    # upwards or alternating; fused pair, alternating; whole step in one kernel
    # (FP32, npc = 2; otherwise the fused pair)
    for fuse, alt, full in ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 1), (1, 1, 2)):
        g = lyo.Lyotropic(LX, LY, npc=npc, precision=precision, **CONFIG0_PHYS)
        g.set_schedule(fuse, xc2, alt, full, xc2)
        g.configure(**CONFIG0_INIT)
        g.step(4)
        mid = [g.get(n) for n in ("ux", "sigmaXY", "MU")]
        g.step(3)
        runs.append((g, mid))
    a, ma = runs[0]
    for b, mb in runs[1:]:
        for x, y in zip(ma, mb):
            assert np.array_equal(x, y)
        for n in STATE + DERIVED:
            x, y = a.get(n), b.get(n)
            assert np.array_equal(x, y), "%s differs: max |d| = %g" % (n, np.abs(x - y).max())


@pytest.mark.parametrize("npc", [1, 2, 3])
@pytest.mark.parametrize("LX,LY", [(128, 128), (45, 67), (5, 3), (2, 60)])
def test_fp64_loop_kernels_bit_exact(require_gpu, LX, LY, npc):
    """The reference's loop structure (one kernel per node loop, derived fields
    stored) gives the oracle's fields bit for bit, and the fused kernels'
    results when the two alternate within one run."""
    g, o, _, _ = pair(LX, LY, npc=npc)
    g.set_loop_kernels(True)
    g.step(1)
    o.step(1)
    assert_bitwise(g, o, STATE)
    assert_bitwise(g, o, DERIVED)
    f = lyo.Lyotropic(LX, LY, npc=npc, precision=lyo.FP64, **CONFIG0_PHYS)
    f.configure(**CONFIG0_INIT)
    f.step(1)
    for on in (False, True, False):
        g.set_loop_kernels(on)
        g.step(2)
        f.step(2)
    for n in STATE + DERIVED:
        x, y = g.get(n), f.get(n)
        assert np.array_equal(x, y), "%s differs: max |d| = %g" % (n, np.abs(x - y).max())


def test_loop_kernels_are_fp64_only(require_gpu):
    g = lyo.Lyotropic(16, 16, npc=2, precision=lyo.FP32, **CONFIG0_PHYS)
    with pytest.raises(lyo.MassLyoError) as e:
        g.set_loop_kernels(True)
    assert e.value.code == 1 and "FP64" in str(e.value)
    g.set_loop_kernels(False)


@pytest.mark.parametrize("precision", [lyo.FP64, lyo.FP32])
@pytest.mark.parametrize("LX,LY,xc", [(128, 128, 32), (45, 67, 7), (5, 3, 1), (8, 60, 3),
                                      (70, 113, 64), (40, 117, 16), (24, 58, 5)])
def test_one_kernel_step_npc1_equals_sweeps(require_gpu, precision, LX, LY, xc):
    """npc = 1: predictor and corrector in one kernel give the two sweeps'
    state and frame fields bit for bit (LX < 8 runs the sweeps)."""
    runs = []
    for full in (0, 1):
        g = lyo.Lyotropic(LX, LY, npc=1, precision=precision, **CONFIG0_PHYS)
        g.set_schedule(one_kernel_step=full, step_chunk_rows=xc)
        g.configure(**CONFIG0_INIT)
        g.step(3)
        mid = [g.get(n) for n in ("ux", "sigmaXY", "MU")]
        g.step(4)
        runs.append((g, mid))
    (a, ma), (b, mb) = runs
    for x, y in zip(ma, mb):
        assert np.array_equal(x, y)
    for n in STATE + DERIVED:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x, y), "%s differs: max |d| = %g" % (n, np.abs(x - y).max())


@pytest.mark.parametrize("precision", [lyo.FP64, lyo.FP32])
@pytest.mark.parametrize("LX,LY,xc", [(24, 700, 9), (30, 500, 37), (16, 250, 8), (40, 244, 8),
                                      (40, 245, 13), (16, 20, 8), (9, 489, 30)])
def test_block_cooperative_step_equals_sweeps(require_gpu, precision, LX, LY, xc):
    """The block-cooperative one-kernel step (four waves per block trading
    their edge values through LDS) gives the separate sweeps' state and frame
    fields bit for bit: grids narrower than a block, exactly one and just over
    one block wide, several blocks with a ragged last one, short chunks."""
    print("block step %dx%d xc=%d precision %d" % (LX, LY, xc, precision), flush=True)
    runs = []
    for fuse, full in ((0, 0), (1, 2)):
        g = lyo.Lyotropic(LX, LY, npc=2, precision=precision, **CONFIG0_PHYS)
        g.set_schedule(fuse_last_pair=fuse, one_kernel_step=full, step_chunk_rows=xc)
        g.configure(**CONFIG0_INIT)
        g.step(4)
        mid = [g.get(n) for n in ("ux", "sigmaXY", "MU")]
        g.step(3)
        runs.append((g, mid))
    (a, ma), (b, mb) = runs
    for x, y in zip(ma, mb):
        assert np.array_equal(x, y)
    for n in STATE + DERIVED:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x, y), "%s differs: max |d| = %g" % (n, np.abs(x - y).max())


def test_fused_pair_timing_kinds(require_gpu):
    # (predictor, separate corrector, fused pair, one-kernel step)
    for npc, full, kinds in ((1, 0, (1, 1, 0, 0)), (1, 1, (0, 0, 0, 1)), (2, 0, (1, 0, 1, 0)),
                             (2, 1, (0, 0, 0, 1)), (2, 2, (0, 0, 0, 1)), (3, 1, (1, 1, 1, 0))):
        g = lyo.Lyotropic(64, 64, npc=npc, precision=lyo.FP32, **CONFIG0_PHYS)
        g.set_schedule(one_kernel_step=full)
        g.configure(**CONFIG0_INIT)
        _, ms = g.time_steps(2)
        assert tuple(int(m > 0) for m in ms) == kinds, (npc, ms)


@pytest.mark.parametrize("precision", [lyo.FP64, lyo.FP32])
def test_plane_padding_does_not_change_results(require_gpu, precision):
    outs = []
    for pad in (0, 4160):
        g = lyo.Lyotropic(61, 83, npc=2, precision=precision, plane_pad=pad, **CONFIG0_PHYS)
        g.configure(**CONFIG0_INIT)
        g.step(4)
        outs.append([g.get(n) for n in STATE + ["sigmaXY"]])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_sums_match_oracle(require_gpu):
    g, o, _, _ = pair(128, 96, npc=1)
    g.step(10)
    o.step(10)
    fs, ps = g.sums()
    s = o.sums()
    # the device sums in a fixed tree; compare with the exactly rounded sum of
    # the (bit-identical) fields, and with the reference's serial sum
    exact_f = math.fsum(o.get("ff").reshape(-1))
    exact_p = math.fsum(o.get("phi"))
    assert fs == pytest.approx(exact_f, rel=4e-15)
    assert ps == pytest.approx(exact_p, rel=4e-15)
    assert fs == pytest.approx(s["fsum"], rel=1e-13)
    assert ps == pytest.approx(s["phisum"], rel=1e-13)
    # fixed tree: repeated calls agree exactly
    assert g.sums() == (fs, ps)
    g.runtime_checks()


def test_fp32_initial_and_100_steps(require_gpu):
    """FP32 (shifted populations, compensated phi) against the FP64 oracle.
    Tolerances: initial state to FP32 rounding; after 100 steps populations
    within 1e-6 relative to rho and Q/phi within 1e-5 absolute."""
    LX, LY = 128, 128
    g, o, _, _ = pair(LX, LY, npc=1, precision=lyo.FP32)
    assert np.allclose(g.get("ff"), o.get("ff"), rtol=0, atol=40 * 1e-7)
    assert np.allclose(g.get("phi"), o.get("phi"), rtol=0, atol=1e-15)  # phi + phic
    g.step(100)
    o.step(100)
    assert np.abs(g.get("ff") - o.get("ff")).max() < 40 * 1e-6
    for n in ("QQxx", "QQyx", "phi"):
        assert np.abs(g.get(n) - o.get(n)).max() < 1e-5, n
    u_g, u_o = g.get("ux"), o.get("ux")
    assert np.abs(u_g - u_o).max() < 1e-3 * np.abs(u_o).max()


def test_fp32_phi_conservation(require_gpu):
    """The compensated update keeps sum(phi) to far better than FP32 rounding."""
    g, _, _, pt = pair(256, 256, npc=2, precision=lyo.FP32)
    p0 = g.sums()[1]
    g.step(200)
    p1 = g.sums()[1]
    # uncompensated FP32 drifts ~ sqrt(N*steps)*ulp(0.5) ~ 1e-3; compensated << that
    assert abs(p1 - p0) < 1e-6 * p0


def test_errors_are_codes_not_crashes(require_gpu):
    with pytest.raises(lyo.MassLyoError) as e:
        lyo.Lyotropic(1, 10, **CONFIG0_PHYS)
    assert e.value.code == 1
    with pytest.raises(lyo.MassLyoError):
        lyo.Lyotropic(10, 10, npc=0, **CONFIG0_PHYS)
    g = lyo.Lyotropic(8, 8, **CONFIG0_PHYS)
    N = 64
    ff = np.tile(po.equilibrium(0, 0, 40.0), (N, 1))
    phi = np.full(N, 0.5)
    phi[5] = np.nan
    g.set_state(ff, np.zeros(N), np.zeros(N), phi)
    with pytest.raises(lyo.MassLyoError) as e:
        g.get("phi")
    assert e.value.code == 3 and "nan" in str(e.value)
    with pytest.raises(lyo.MassLyoError):
        g.configure(config=99)


def test_precision_report_small(require_gpu):
    """FP32 against FP64 over 400 steps at 128^2, npc=2: per-field relative
    errors stay small and the compensated phi drifts less than FP64 rounding
    allows for an uncompensated FP32 sum."""
    from mass_hd_amd.reports import precision_report
    r = precision_report(size=128, steps=400, npc=2, chunks=2)
    last = r["series"][-1]
    assert last["ff"]["max_rel"] < 1e-5
    assert last["QQxx"]["max_rel"] < 1e-3 and last["phi"]["max_rel"] < 1e-3
    assert abs(last["phi_drift_fp32"]) < 1e-9
    assert abs(last["phi_drift_fp64"]) < 1e-12
