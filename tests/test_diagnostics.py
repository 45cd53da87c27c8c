"""Diagnostics: rms velocity, phi statistics/histogram and nematic defects.

CPU: the numpy restatement of the reference's post-processing
(oracle/diag_oracle.py) is pinned by analytic defect fields of known charge and
by an independent component count (scipy.ndimage.label plus periodic merging).
GPU: mass_lyo_diagnostics equals the restatement on the same fields.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import diag_oracle as dg


def pair_field(LX, LY, cores):
    """Director angle sum_i s_i/2 * atan2(y - y_i, x - x_i): a defect of charge
    s_i/2 at each core (cores off the lattice)."""
    x, y = np.meshgrid(np.arange(LX), np.arange(LY), indexing="ij")
    th = np.zeros((LX, LY))
    for (cx, cy, s) in cores:
        th += 0.5 * s * np.arctan2(y - cy, x - cx)
    return (np.cos(2 * th)).reshape(-1), (np.sin(2 * th)).reshape(-1)


def periodic_components(mask):
    """Independent count of 4-connected components on a torus."""
    lab, n = ndimage.label(mask)
    parent = list(range(n + 1))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    LX, LY = mask.shape
    for i in range(LX):
        a, b = lab[i, 0], lab[i, LY - 1]
        if a and b:
            parent[find(a)] = find(b)
    for j in range(LY):
        a, b = lab[0, j], lab[LX - 1, j]
        if a and b:
            parent[find(a)] = find(b)
    return len({find(i) for i in range(1, n + 1)})


def test_charge_of_analytic_pair():
    LX, LY = 40, 40
    qxx, qyx = pair_field(LX, LY, [(15.5, 20.5, +1), (25.5, 20.5, -1)])
    w = dg.charge(qxx, qyx, LX, LY)
    assert np.isclose(w.max(), 0.5) and np.isclose(w.min(), -0.5)
    d = dg.defects(qxx, qyx, LX, LY)
    assert sorted(c for (_, _, c) in d) == [-0.5, 0.5]
    pos = [p for p in d if p[2] > 0][0]
    assert abs(pos[0] - 15.5) <= 0.5 and abs(pos[1] - 20.5) <= 0.5


def test_charge_sign_convention():
    """theta = +atan2/2 winds +pi around the core (charge +1/2)."""
    qxx, qyx = pair_field(30, 30, [(10.5, 14.5, +1), (19.5, 14.5, -1)])
    w = dg.charge(qxx, qyx, 30, 30)
    assert w[10, 14] == pytest.approx(0.5) and w[19, 14] == pytest.approx(-0.5)


@pytest.mark.parametrize("seed", range(6))
def test_defect_count_matches_independent_components(seed):
    rng = np.random.default_rng(seed)
    LX, LY = 17, 23
    th = rng.uniform(0, np.pi, (LX, LY))
    th = ndimage.uniform_filter(th, 2, mode="wrap")  # some structure, many defects
    qxx, qyx = np.cos(2 * th).reshape(-1), np.sin(2 * th).reshape(-1)
    w = dg.charge(qxx, qyx, LX, LY)
    n_ref = periodic_components(w > .4) + periodic_components(w < -.4)
    assert len(dg.defects(qxx, qyx, LX, LY)) == n_ref


def test_histogram_bins():
    phi = np.array([-0.1, 0.0, 1 / 64 - 1e-12, 1 / 64, 0.5, 0.999999, 1.0, 2.0, np.nan])
    h = dg.phi_histogram(phi)
    assert h[0] == 2 and h[-1] == 2 and h[1] == 2 and h[2] == 1 and h[33] == 1 and h[64] == 1
    assert h.sum() == len(phi)


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_diagnostics_match_restatement(require_gpu, precision):
    from mass_hd_amd import lyo
    from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS
    LX, LY = 96, 80
    m = lyo.Lyotropic(LX, LY, npc=1, precision=precision, **CONFIG0_PHYS)
    m.configure(**CONFIG0_INIT)
    m.step(30)
    d = m.diagnostics(charge=True)
    ff, qxx, qyx, phi = m.get("ff"), m.get("QQxx"), m.get("QQyx"), m.get("phi")
    vx, vy = dg.velocity(ff, LX, LY)
    u2 = vx ** 2 + vy ** 2
    assert d["u_rms"] == pytest.approx(math.sqrt(u2.mean()), rel=1e-9)
    assert d["u_mean"] == pytest.approx(np.sqrt(u2).mean(), rel=1e-9)
    assert d["phi_mean"] == pytest.approx(phi.mean(), rel=1e-12)
    assert d["phi_min"] == pytest.approx(phi.min(), rel=1e-12)
    assert d["phi_max"] == pytest.approx(phi.max(), rel=1e-12)
    assert np.array_equal(d["phi_hist"], dg.phi_histogram(phi))
    w = dg.charge(qxx, qyx, LX, LY)
    assert np.allclose(d["charge"].reshape(LX, LY), w, atol=1e-12, equal_nan=True)
    ref = dg.defects(qxx, qyx, LX, LY)
    assert d["defects_pos"] == sum(1 for p in ref if p[2] > 0)
    assert d["defects_neg"] == sum(1 for p in ref if p[2] < 0)
    assert d["defect_density"] == pytest.approx(len(ref) / (LX * LY))
    assert d["defects_pos"] + d["defects_neg"] > 0


@pytest.mark.gpu
def test_gpu_defects_of_analytic_pairs(require_gpu):
    from mass_hd_amd import lyo
    from oracle import pyoracle as po
    from tests.conftest import CONFIG0_PHYS
    LX, LY = 64, 48
    cores = [(10.5, 10.5, 1), (20.5, 10.5, -1), (40.5, 30.5, 1), (50.5, 30.5, -1),
             (30.5, 40.5, -1), (30.5, 24.5, 1)]
    qxx, qyx = pair_field(LX, LY, cores)
    m = lyo.Lyotropic(LX, LY, **CONFIG0_PHYS)
    N = LX * LY
    m.set_state(np.tile(po.equilibrium(0, 0, 40.0), (N, 1)), qxx, qyx, np.full(N, 0.5))
    d = m.diagnostics()
    assert (d["defects_pos"], d["defects_neg"]) == (3, 3)
    assert d["u_rms"] == 0.0
