"""go-or-grow on the GPU against the CPU oracle.

The go-or-grow kernels (mass_hd_amd/csrc/lyo_growth.hpp) follow the
reference's loop structure in FP64 and draw the masks from the same mt19937
stream on the host, so every field -- the Lyotropic state, the derived frame
fields, m, chi and the division/death masks -- must equal the oracle's bit for
bit after configure (dry relaxation, hydrodynamic reset, phenotype set-up) and
after steps that redraw the masks, on periodic grids (ghost layer with
periodic ghosts) and walled ones.  With conserve_phi (no division, no death)
phi's sum enters every step; the device sums it in a fixed tree where the
reference sums serially, so those runs are compared at 1e-12.
"""
import numpy as np
import pytest

from mass_hd_amd import lyo
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

PHYS = dict(GammaQ=0.3, GammaP=5e-2, xi=0.3, tauNem=1.0, tauIso=1.0, rho=40.0, friction=0.0,
            AA=0.03, CC=0.1, LL=0.002, KK=0.06, zeta=0.02, zetaI=0.0, conc=1.0)
# the proliferation cases of the reference's runcards, scaled down
GROWTH = dict(B=0.1, Snem=1.0, relax_steps=20, relax_phi=1, relax_Q=1, relax_dt=1.0,
              chi_config="noise", chi0=1.0, chi_noise=0.0, chi_length=0.0, Dchi=0.0,
              Achi=0.0, Ochi=0.0, phiswitch=0.0, growTogether=0, beta=0.0, death_rate=0.0,
              death_time=1.0, death_radius=0.0, division_rate=0.1, division_time=2.0,
              division_radius=3.0, alpha=0.001, phi_critical=1.2)
INIT = dict(config="circle", seed=1001, radius=10.0, level=20.0, noise=1.0, angle=0.0)
STATE = ["ff", "QQxx", "QQyx", "phi", "m", "chi"]
DERIVED = ["n", "ux", "uy", "sigmaXX", "sigmaXY", "sigmaYX", "sigmaYY", "sigma_bulk",
           "sigma_elastic_xx", "sigma_elastic_yx", "sigma_phase_field_xx",
           "sigma_phase_field_yx", "sigma_active_xx", "sigma_active_yx", "MU", "HHxx", "HHyx"]


def pair(LX, LY, bc=0, npc=1, phys=PHYS, growth=None, init=INIT):
    g = dict(GROWTH)
    g.update(growth or {})
    a = lyo.Lyotropic(LX, LY, npc=npc, precision=lyo.FP64, bc=bc, **phys)
    a.set_growth(**g)
    o = po.Oracle(LX, LY, npc=npc, bc=bc, **phys)
    o.set_growth(**g)
    ta = a.configure(**init)
    to = o.configure(**init)
    return a, o, ta, to


def check(a, o, names, what=""):
    for n in names:
        x, y = a.get(n), o.get(n)
        assert np.array_equal(x, y), "%s %s differs: max |d| = %g" % (what, n, np.abs(x - y).max())
    da, ea = a.get("division_mask"), a.get("death_mask")
    do, eo = o.masks()
    assert np.array_equal(da, do.astype(float)) and np.array_equal(ea, eo.astype(float)), what


@pytest.mark.parametrize("bc", [0, 2, 4])
@pytest.mark.parametrize("npc", [1, 2])
@pytest.mark.parametrize("LX,LY", [(40, 36), (7, 5)])
def test_growth_bit_exact(require_gpu, bc, npc, LX, LY):
    print("go-or-grow bc=%d npc=%d %dx%d" % (bc, npc, LX, LY), flush=True)
    a, o, ta, to = pair(LX, LY, bc, npc)
    assert ta == to  # ftot of the reset, ptot of the relaxed phi
    check(a, o, STATE, "configure")
    for s in range(3):
        a.step(4)
        o.step(4)
        check(a, o, STATE + DERIVED, "after %d steps" % (4 * s + 4))
    if LX > 10:
        assert a.get("division_mask").any()  # the masks were drawn and used
        assert a.get("phi").sum() > ta[1]     # division grew the colony


@pytest.mark.parametrize("growth", [
    dict(chi_noise=0.01, chi_length=2.0, chi0=0.6, Dchi=0.05, Achi=0.02, Ochi=0.01,
         phiswitch=0.5, growTogether=1),
    dict(chi_config="front", chi_length=4.0, beta=0.0005, death_rate=0.05, death_time=3.0,
         death_radius=2.0),
])
@pytest.mark.parametrize("bc", [0, 3])
def test_growth_phenotype_terms_bit_exact(require_gpu, growth, bc):
    a, o, ta, to = pair(48, 40, bc, npc=2, growth=growth)
    assert ta == to
    check(a, o, STATE, "configure")
    a.step(7)
    o.step(7)
    check(a, o, STATE + DERIVED, "7 steps")


def test_growth_conserve_phi_matches_to_rounding(require_gpu):
    # no division or death: conserve_phi adds (countphi - totalphi)/N to Dp
    a, o, ta, to = pair(40, 36, 0, npc=1, growth=dict(alpha=0.0, division_rate=0.0))
    assert ta == to
    a.step(10)
    o.step(10)
    for n in ["ff", "QQxx", "QQyx", "phi", "m"]:
        x, y = a.get(n), o.get(n)
        assert np.allclose(x, y, rtol=1e-12, atol=1e-14), n
    # chi = m/phi amplifies those differences where phi is small
    phi = o.get("phi")
    big = phi > 1e-3
    assert np.allclose(a.get("chi")[big], o.get("chi")[big], rtol=1e-8, atol=0)


def test_growth_512_bounds_and_repeats(require_gpu):
    print("go-or-grow 512^2", flush=True)
    runs = []
    for _ in range(2):
        a = lyo.Lyotropic(512, 512, npc=1, precision=lyo.FP64, **PHYS)
        a.set_growth(**dict(GROWTH, relax_steps=50))
        a.configure(**dict(INIT, radius=120.0))
        a.step(30)
        phi, m = a.get("phi"), a.get("m")
        assert np.all(m >= 0) and np.all(m <= np.maximum(phi, 0))  # ProjectM's bounds
        assert np.isfinite(a.get("ff")).all()
        runs.append([a.get(n) for n in STATE])
    for x, y in zip(*runs):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("division_only", [0, 1])
@pytest.mark.parametrize("npc", [1, 2])
def test_growth_fused_sweeps_equal_loops(require_gpu, npc, division_only):
    """Periodic go-or-grow (and lyotropic-with-division) as fused tile sweeps
    against the reference's node loops on the same run: bitwise every state
    and frame field, m, chi and the masks, on a grid of partial tiles
    (130 x 70 against 10 x 62 tiles), over steps with mask redraws."""
    print("go-or-grow fused vs loops npc=%d division_only=%d" % (npc, division_only), flush=True)
    g = dict(GROWTH, relax_steps=10, division_only=division_only)
    runs = []
    for loops in (False, True):
        a = lyo.Lyotropic(130, 70, npc=npc, precision=lyo.FP64, **PHYS)
        a.set_growth(**g)
        a.set_loop_kernels(loops)
        a.configure(**dict(INIT, radius=25.0, level=35.0))
        a.step(7)
        names = (["ff", "QQxx", "QQyx", "phi"] + ([] if division_only else ["m", "chi"]) +
                 DERIVED + ["division_mask", "death_mask"])
        runs.append({n: a.get(n) for n in names})
        a.close()
    for n, x in runs[0].items():
        assert np.array_equal(x, runs[1][n]), n
    assert runs[0]["division_mask"].any()


def test_growth_fused_sweeps_equal_loops_large(require_gpu):
    """The fused sweeps against the node loops on a large grid of odd sizes
    (1003 x 997: partial tiles in both directions), with masks drawn."""
    print("go-or-grow fused vs loops 1003x997", flush=True)
    g = dict(GROWTH, relax_steps=10)
    runs = []
    for loops in (False, True):
        a = lyo.Lyotropic(1003, 997, npc=1, precision=lyo.FP64, **PHYS)
        a.set_growth(**g)
        a.set_loop_kernels(loops)
        a.configure(**dict(INIT, radius=200.0, level=500.0))
        a.step(6)
        runs.append({n: a.get(n) for n in ["ff", "QQxx", "phi", "m", "chi", "sigmaXY",
                                            "division_mask"]})
        a.close()
    for n, x in runs[0].items():
        assert np.array_equal(x, runs[1][n]), n
    assert runs[0]["division_mask"].any()


def test_growth_count_only_redraws_keep_the_draw_order(require_gpu):
    """Redraws with rate zero accept nothing and only consume one number per
    listed node; the engine reads their list lengths back without waiting and
    consumes the numbers at the next redraw that accepts.  Death redrawn every
    step at rate zero and division every 101st step: 100 such redraws between
    two accepting ones (past the 64 whose lengths may be in flight), then the
    division draws must see the oracle's numbers -- masks and state bit for
    bit."""
    print("go-or-grow count-only redraws", flush=True)
    g = dict(division_time=100.0, division_rate=0.5, death_time=0.0, death_rate=0.0,
             beta=0.0, relax_steps=5)
    a, o, ta, to = pair(40, 36, growth=g)
    for s in range(3):
        a.step(70)
        o.step(70)
        check(a, o, STATE, "after %d steps" % (70 * s + 70))
    assert a.get("division_mask").any()


def test_growth_refusals(require_gpu):
    f32 = lyo.Lyotropic(16, 16, npc=1, precision=lyo.FP32, **PHYS)
    with pytest.raises(lyo.MassLyoError):
        f32.set_growth(**GROWTH)
    plain = lyo.Lyotropic(16, 16, npc=1, precision=lyo.FP64, **PHYS)
    plain.configure(**INIT)
    with pytest.raises(lyo.MassLyoError):
        plain.get("m")
    a = lyo.Lyotropic(16, 16, npc=1, precision=lyo.FP64, **PHYS)
    with pytest.raises(lyo.MassLyoError):
        a.set_growth(**dict(GROWTH, chi_noise=-1.0))
    a.set_growth(**GROWTH)
    a.configure(**INIT)
    a.set_loop_kernels(False)  # periodic: the fused sweeps (the default)
    w = lyo.Lyotropic(16, 16, npc=1, precision=lyo.FP64, bc=2, **PHYS)
    w.set_growth(**GROWTH)
    with pytest.raises(lyo.MassLyoError):
        w.set_loop_kernels(False)  # walled go-or-grow runs the node loops only
    for call in (a.diagnostics, a.checkpoint):
        with pytest.raises(lyo.MassLyoError):
            call()
    a.step(2)
    assert np.isfinite(a.get("m")).all()


@pytest.mark.parametrize("bc", [0, 2])
@pytest.mark.parametrize("npc", [1, 2])
def test_division_only_bit_exact(require_gpu, bc, npc):
    # lyotropic-with-division: Lyotropic with the masks' phi growth and death
    div = dict(division_only=1, alpha=0.002, beta=0.001, division_rate=0.1, division_time=2.0,
               division_radius=3.0, death_rate=0.05, death_time=1.0, death_radius=2.0,
               phi_critical=1.2)
    a, o, ta, to = pair(40, 36, bc, npc, phys=dict(PHYS, zeta=0.03), growth=div)
    assert ta == to
    for s in range(3):
        a.step(3)
        o.step(3)
        check(a, o, ["ff", "QQxx", "QQyx", "phi"] + DERIVED, "after %d steps" % (3 * s + 3))
    assert a.get("division_mask").any() and a.get("death_mask").any()
    with pytest.raises(lyo.MassLyoError):
        a.get("m")  # no phenotype in this model


def test_growth_reconfigure_restarts_the_draws(require_gpu):
    # configure starts the run afresh (generator reseeded, masks cleared and
    # due at the first step): a run configured again after redrawing its
    # masks must match the oracle after the same calls, and a fresh oracle
    a, o, ta, to = pair(40, 36, 0, npc=1)
    a.step(7)
    o.step(7)
    assert a.get("division_mask").any()
    assert a.configure(**INIT) == o.configure(**INIT)
    o2 = po.Oracle(40, 36, npc=1, bc=0, **PHYS)
    o2.set_growth(**GROWTH)
    o2.configure(**INIT)
    for s in range(2):
        a.step(5)
        o.step(5)
        o2.step(5)
        check(a, o, STATE, "after reconfigure, %d steps" % (5 * s + 5))
        check(a, o2, STATE, "against a fresh run, %d steps" % (5 * s + 5))


def test_growth_long_node_lists_bit_exact(require_gpu):
    # about 2,000 nodes with phi >= 1/2 over four selection blocks: the first
    # redraw reads back more of the device's node list than its first guess
    a, o, ta, to = pair(64, 64, 0, npc=1, init=dict(INIT, radius=25.0))
    assert ta == to
    assert (o.get("phi") >= 0.5).sum() > 1024
    for s in range(2):
        a.step(4)
        o.step(4)
        check(a, o, STATE, "after %d steps" % (4 * s + 4))


def test_growth_set_state_is_refused(require_gpu):
    # m, chi, the masks and the draws ahead cannot be set through set_state
    a = lyo.Lyotropic(12, 10, npc=1, precision=lyo.FP64, **PHYS)
    a.set_growth(**GROWTH)
    a.configure(**INIT)
    with pytest.raises(lyo.MassLyoError):
        a.set_state(a.get("ff"), a.get("QQxx"), a.get("QQyx"), a.get("phi"))
    a.close()
