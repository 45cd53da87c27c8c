"""Pins the oracle's go-or-grow model (src/models/go-or-grow.cpp,
src/models/lyotropic-with-division.cpp): libstdc++'s normal draws bit for
bit, the division masks against a brute-force dilation and their refresh
period, the free-energy derivatives (mu = f'(phi), H = -df/dQ), the dry
relaxation as a descent that keeps phi, the phenotype density reducing to
phi when chi = 1, the phenotype diffusion and switching terms against the
predictor-corrector recurrences of their linear / ODE forms, and the growth
source's per-step recurrence.
"""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po

PHYS = dict(GammaQ=0.3, GammaP=5e-2, xi=0.3, tauNem=1.0, tauIso=1.0, rho=40.0, friction=0.0,
            AA=0.03, CC=0.1, LL=0.002, KK=0.06, zeta=0.0, zetaI=0.0, conc=1.0)
NO_GROWTH = dict(alpha=0.0, beta=0.0, division_rate=0.0, division_time=1.0, division_radius=0.0,
                 death_rate=0.0, death_time=1.0, death_radius=0.0, phi_critical=1.2)


def uniform_run(LX, LY, phi=1.0, npc=1, phys=None, **growth):
    """A go-or-grow run at rest with phi uniform and Q = 0."""
    p = dict(PHYS)
    p.update(phys or {})
    o = po.Oracle(LX, LY, npc=npc, **dict(p, conc=phi))
    g = dict(NO_GROWTH, chi0=1.0)
    g.update(growth)
    o.set_growth(**g)
    o.configure(config="random", seed=1, qamp=0.0, phinoise=0.0, noise=0.0)
    return o


_NORMAL = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char** argv) {
  std::mt19937 gen(std::strtoul(argv[1], 0, 10));
  for (int i = 0; i < std::atoi(argv[2]); ++i)
    std::printf("%a\n", std::normal_distribution<>(0., 1.)(gen));
  return 0;
}
"""


@pytest.mark.parametrize("seed", [1, 1001, 4294967295])
def test_random_normal_matches_libstdcxx(seed):
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "n.cpp"), os.path.join(d, "n")
        open(src, "w").write(_NORMAL)
        subprocess.check_call(["g++", "-O2", "-o", exe, src])
        out = subprocess.check_output([exe, str(seed), "3000"], text=True).split()
    ref = np.array([float.fromhex(t) for t in out])
    assert np.array_equal(po.random_normal(seed, 3000), ref)
    assert abs(ref.mean()) < 0.1 and abs(ref.std() - 1) < 0.05


def test_division_masks_are_dilated_centres_and_refresh_on_schedule():
    # rate 2 accepts every node with phi >= .5: the mask is that set dilated
    # by the periodic disk of squared radius ceil(r^2)
    LX, LY, r = 40, 30, 2.12  # r^2 = 4.49: ceil admits d^2 = 5
    o = po.Oracle(LX, LY, npc=1, **PHYS)
    o.set_growth(**dict(NO_GROWTH, division_rate=2.0, division_time=2.0, division_radius=r,
                        alpha=1e-4, phi_critical=1e9, chi0=1.0))
    o.configure(config="circle", seed=1, radius=6.0, noise=0.0)
    src = o.get("phi") >= .5
    o.step(1)
    div, death = o.masks()
    x, y = np.meshgrid(np.arange(LX), np.arange(LY), indexing="ij")
    x, y = x.reshape(-1), y.reshape(-1)
    want = np.zeros(LX * LY, bool)
    for c in np.nonzero(src)[0]:
        dx = np.abs(x - x[c]) % LX
        dy = np.abs(y - y[c]) % LY
        d2 = np.minimum(dx, LX - dx) ** 2 + np.minimum(dy, LY - dy) ** 2
        want |= d2 <= np.ceil(r * r)
    assert np.array_equal(div.astype(bool), want) and not death.any()
    # division_time = 2: the mask is kept for two more steps, then redrawn
    o2 = po.Oracle(LX, LY, npc=1, **PHYS)
    o2.set_growth(**dict(NO_GROWTH, division_rate=0.3, division_time=2.0, division_radius=1.0,
                         alpha=1e-4, phi_critical=1e9, chi0=1.0))
    o2.configure(config="circle", seed=5, radius=8.0, noise=0.0)
    seen = []
    for _ in range(7):
        o2.step(1)
        seen.append(o2.masks()[0].copy())
    assert np.array_equal(seen[0], seen[1]) and np.array_equal(seen[1], seen[2])
    assert not np.array_equal(seen[2], seen[3])
    assert np.array_equal(seen[3], seen[5]) and not np.array_equal(seen[5], seen[6])


def derived_at(phi, qxx, qyx, **growth):
    o = uniform_run(4, 4, phi=phi, **growth)
    N = 16
    o.set_field("QQxx", np.full(N, qxx))
    o.set_field("QQyx", np.full(N, qyx))
    o.update_quantities()
    return {k: o.get(k)[0] for k in ("MU", "HHxx", "HHyx", "sigma_bulk")}


@pytest.mark.parametrize("phi", [0.3, 1.1, 1.35])
def test_chemical_potential_and_molecular_field_are_free_energy_derivatives(phi):
    # sigma_bulk = f - mu*phi with f = AA/2 p^2 (1-p)^2 + B/2 (p - pc)_+^2
    # + CC/2 (Snem - |Q|^2)^2; mu = df/dphi, H = -df/dQ (go-or-grow.cpp:257-276)
    g = dict(B=0.1, Snem=0.8, phi_critical=1.2)
    qxx, qyx, h = 0.3, -0.2, 1e-5

    def f(p, a, b):
        d = derived_at(p, a, b, **g)
        return d["sigma_bulk"] + d["MU"] * p

    d0 = derived_at(phi, qxx, qyx, **g)
    mu_fd = (f(phi + h, qxx, qyx) - f(phi - h, qxx, qyx)) / (2 * h)
    hxx_fd = -(f(phi, qxx + h, qyx) - f(phi, qxx - h, qyx)) / (2 * h)
    hyx_fd = -(f(phi, qxx, qyx + h) - f(phi, qxx, qyx - h)) / (2 * h)
    assert d0["MU"] == pytest.approx(mu_fd, rel=1e-6, abs=1e-9)
    assert d0["HHxx"] == pytest.approx(hxx_fd, rel=1e-6, abs=1e-9)
    assert d0["HHyx"] == pytest.approx(hyx_fd, rel=1e-6, abs=1e-9)


def free_energy(o, g):
    p, qx, qy = o.get("phi"), o.get("QQxx"), o.get("QQyx")
    LX, LY = o.LX, o.LY

    def grad2(a):
        a = a.reshape(LX, LY)
        return ((np.roll(a, -1, 0) - a) ** 2 + (np.roll(a, -1, 1) - a) ** 2).sum()
    dp = np.maximum(p - g["phi_critical"], 0)
    bulk = (.5 * PHYS["AA"] * p ** 2 * (1 - p) ** 2 + .5 * g["B"] * dp ** 2
            + .5 * PHYS["CC"] * (g["Snem"] - qx ** 2 - qy ** 2) ** 2).sum()
    return bulk + .5 * PHYS["KK"] * grad2(p) + .5 * PHYS["LL"] * (grad2(qx) + grad2(qy))


def test_dry_relaxation_descends_and_keeps_phi():
    g = dict(NO_GROWTH, B=0.1, Snem=1.0, chi0=1.0, relax_phi=1, relax_Q=1, relax_dt=1.0)
    energies, sums = [], []
    for steps in (0, 5, 20, 80):
        o = po.Oracle(48, 40, npc=1, **PHYS)
        o.set_growth(**dict(g, relax_steps=steps))
        o.configure(config="circle", seed=1001, radius=12.0, noise=1.0)
        energies.append(free_energy(o, g))
        sums.append(o.get("phi").sum())
        # ResetHydrodynamics: rest equilibrium everywhere afterwards
        assert np.array_equal(o.get("ff"), np.tile(po.equilibrium(0., 0., 40.), (48 * 40, 1)))
    assert all(b < a for a, b in zip(energies, energies[1:]))
    assert np.allclose(sums, sums[0], rtol=1e-12)


def test_relaxation_keeps_a_state_at_the_bulk_minima():
    # KK = LL = 0: phi in {0, 1} and |Q|^2 = Snem (or Q = 0) make mu and H
    # vanish, so the descent (phi += dt GammaP lap mu, Q += dt GammaQ H;
    # go-or-grow.cpp:80-108) leaves the disc untouched, where any diffusive
    # update would smear its edge
    o = po.Oracle(32, 32, npc=1, **dict(PHYS, KK=0.0, LL=0.0))
    o.set_growth(**dict(NO_GROWTH, chi0=1.0, Snem=1.0, relax_steps=30, relax_phi=1, relax_Q=1))
    o.configure(config="circle", seed=1, radius=9.0, angle=0.0, noise=0.0)
    ref = po.Oracle(32, 32, npc=1, **PHYS)
    ref.configure(config="circle", seed=1, radius=9.0, angle=0.0, noise=0.0)
    for n in ("phi", "QQxx", "QQyx"):
        assert np.array_equal(o.get(n), ref.get(n)), n


def test_phenotype_density_is_phi_when_chi_is_one():
    # chi = 1 and no growth: the m equation's fluxes become phi's (GammaP
    # lap mu, div(u m); go-or-grow.cpp:322-339), so m tracks phi (the growth
    # source enters the two differently: test_growth_source_recurrence)
    o = po.Oracle(40, 40, npc=2, **dict(PHYS, zeta=0.01))
    o.set_growth(**dict(NO_GROWTH, chi0=1.0))
    o.configure(config="random", seed=3, angle=0.0, noise=1.0, qamp=0.5, phinoise=0.05)
    o.step(40)
    phi, m = o.get("phi"), o.get("m")
    assert phi.min() > 0
    assert np.allclose(m, phi, rtol=0, atol=1e-10)
    assert np.abs(o.get("ux")).max() > 1e-6


def set_chi(o, chi):
    o.set_field("m", o.get("phi") * chi)
    o.set_field("chi", chi)


@pytest.mark.parametrize("npc", [1, 2])
def test_phenotype_diffusion_mode_decays_by_the_step_recurrence(npc):
    # phi = 1, no flow: dm = Dchi lap(chi) (go-or-grow.cpp:334-339); a
    # Fourier mode has lam = -4 Dchi sin^2(pi n / L) per sweep and the
    # predictor-corrector step multiplies it by 1 + lam + lam^2/2 (npc 1) or
    # 1 + lam + lam^2/2 + lam^3/4 (npc 2)
    L, n, Dchi = 32, 3, 0.1
    o = uniform_run(L, 4, phi=1.0, npc=npc, Dchi=Dchi)
    x = np.arange(L * 4) // 4
    chi = 0.5 + 0.01 * np.cos(2 * np.pi * n * x / L)
    set_chi(o, chi)
    lam = -4 * Dchi * np.sin(np.pi * n / L) ** 2
    G = 1 + lam + lam ** 2 / 2 + (lam ** 3 / 4 if npc == 2 else 0)
    o.step(10)
    got = o.get("chi") - 0.5
    assert np.allclose(got, 0.01 * G ** 10 * np.cos(2 * np.pi * n * x / L), rtol=0, atol=1e-13)


def test_phenotype_switching_is_a_heun_step_of_its_ode():
    # uniform chi: dchi/dt = -phi V'(chi) with V' = 2 Achi chi (1-chi)(1-2chi)
    # + Ochi (phi - phiswitch) (go-or-grow.cpp:340-344); npc = 1 is Heun
    A, O, ps = 0.05, 0.02, 0.4
    o = uniform_run(8, 8, phi=1.0, npc=1, Achi=A, Ochi=O, phiswitch=ps)
    chi = 0.3
    set_chi(o, np.full(64, chi))
    rhs = lambda c: -(2 * A * c * (1 - c) * (1 - 2 * c) + O * (1.0 - ps))
    for _ in range(5):
        c1 = chi + rhs(chi)
        chi = chi + .5 * rhs(chi) + .5 * rhs(c1)
    o.step(5)
    assert np.allclose(o.get("chi"), chi, rtol=1e-14, atol=0)


def test_growth_source_recurrence():
    # every node divides (rate 2, refreshed each step): R = alpha chi phi
    # enters phi whole in each sweep (phi_tmp += R) but m through the
    # predictor-corrector increment (go-or-grow.cpp:345-356)
    a = 0.01
    o = uniform_run(6, 6, phi=1.0, npc=1, alpha=a, division_rate=2.0, division_time=0.0,
                    division_radius=0.0, chi0=0.5, phi_critical=1e9)
    phi, m = 1.0, 0.5
    for _ in range(4):
        R0 = (m / phi) * phi * a
        phi1, m1 = phi + R0, m + R0
        R1 = (m1 / phi1) * phi1 * a
        m = m + .5 * R0 + .5 * R1
        phi = phi + R1
    o.step(4)
    assert np.allclose(o.get("phi"), phi, rtol=1e-14, atol=0)
    assert np.allclose(o.get("m"), m, rtol=1e-14, atol=0)
    assert np.allclose(o.get("chi"), m / phi, rtol=1e-14, atol=0)


def test_phenotype_noise_has_the_requested_mean_and_variance():
    for length in (0.0, 2.0):
        o = po.Oracle(64, 64, npc=1, **PHYS)
        o.set_growth(**dict(NO_GROWTH, chi0=0.5, chi_noise=0.004, chi_length=length))
        o.configure(config="random", seed=9, qamp=0.0, phinoise=0.0)
        chi = o.get("chi")
        assert chi.mean() == pytest.approx(0.5, abs=1e-12)
        assert chi.var() == pytest.approx(0.004, rel=1e-9)
        c = chi.reshape(64, 64) - 0.5
        corr = (c * np.roll(c, 1, 0)).mean() / c.var()
        assert (corr > 0.5) if length else (abs(corr) < 0.05)


def test_phenotype_front():
    o = po.Oracle(30, 6, npc=1, **PHYS)
    o.set_growth(**dict(NO_GROWTH, chi_config="front", chi_length=3.0))
    o.configure(config="random", seed=1, qamp=0.0, phinoise=0.0)
    # chi = .5 (1 - tanh((x - c)/length)), c = (LX-1)/2 (go-or-grow.cpp:183-194),
    # with the C library's tanh (math.tanh) bit for bit
    want = [.5 * (1. - math.tanh((k // 6 - 14.5) / 3.0)) for k in range(30 * 6)]
    assert np.array_equal(o.get("chi"), np.array(want))


def test_division_only_is_lyotropic_without_patches():
    # lyotropic-with-division (lyotropic-with-division.cpp:110-126) adds
    # phi*(alpha*division - beta*death) to Lyotropic's phi update: with no
    # patch drawn it is Lyotropic bit for bit
    from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS
    a = po.Oracle(24, 20, npc=2, **CONFIG0_PHYS)
    a.set_growth(**dict(NO_GROWTH, division_only=1, alpha=1e-3, division_rate=1e-300,
                        division_time=2.0, division_radius=2.0))
    a.configure(**CONFIG0_INIT)
    b = po.Oracle(24, 20, npc=2, **CONFIG0_PHYS)
    b.configure(**CONFIG0_INIT)
    a.step(15)
    b.step(15)
    for n in ("ff", "QQxx", "QQyx", "phi", "sigmaXY", "MU"):
        assert np.array_equal(a.get(n), b.get(n)), n
    assert not a.masks()[0].any()


def test_division_only_growth_recurrence():
    # every node divides: each sweep adds alpha*phi of the phi it reads, so
    # one npc = 1 step maps phi to phi (1 + alpha + alpha^2)
    a = 0.01
    o = po.Oracle(6, 6, npc=1, **PHYS)
    o.set_growth(**dict(NO_GROWTH, division_only=1, alpha=a, division_rate=2.0,
                        division_time=0.0, division_radius=0.0, phi_critical=1e9))
    o.configure(config="random", seed=1, qamp=0.0, phinoise=0.0, noise=0.0)
    o.step(5)
    assert np.allclose(o.get("phi"), (1 + a + a * a) ** 5, rtol=1e-14, atol=0)
