"""Pins the oracle's walled grids (bc 1-4: the ghost layer of
src/geometry.cpp:140-183 and the boundary conditions of
src/models/lyotropic.cpp:388-498, src/fields.hpp:619-918).

The reference ships no vectors for these paths either, so the checks are
physical facts a mistake in a wall or corner rule would break: exact
conservation of mass and phi, the rest state as a fixed point, tangential
momentum kept by free-slip walls, the viscous decay rate of the lowest shear
mode between walls half a node outside the domain, the box's and channel's
mirror, transpose and translation symmetries, and the wall-specific initial
configurations against their closed forms.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.conftest import CONFIG0_INIT, CONFIG0_PHYS

ZERO_PHYS = dict(GammaQ=0.0, GammaP=0.0, xi=0.0, tauNem=1.0, tauIso=1.0, rho=40.0,
                 friction=0.0, AA=0.0, CC=0.0, LL=0.0, KK=0.0, zeta=0.0, zetaI=0.0, conc=0.0)
ACTIVE = dict(CONFIG0_PHYS, zeta=0.05)
C = np.array([(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (-1, 1), (1, -1)])
BCS = [1, 2, 3, 4]


def vel(ff):
    n = ff.sum(1)
    return (ff @ C[:, 0]) / n, (ff @ C[:, 1]) / n


def layer_xy(LX, LY):
    """(x, y) on the (LX+2)*(LY+2) grid of every node index, in the index
    order the ghost layer uses: domain, walls Left/Right/Front/Back, corners
    LeftFront, RightFront, LeftBack, RightBack."""
    xy = [(x + 1, y + 1) for x in range(LX) for y in range(LY)]
    xy += [(0, y + 1) for y in range(LY)] + [(LX + 1, y + 1) for y in range(LY)]
    xy += [(x + 1, 0) for x in range(LX)] + [(x + 1, LY + 1) for x in range(LX)]
    xy += [(0, 0), (LX + 1, 0), (0, LY + 1), (LX + 1, LY + 1)]
    return np.array(xy)


@pytest.mark.parametrize("LX,LY", [(2, 2), (5, 3), (7, 9)])
def test_layer_table_is_the_padded_torus(LX, LY):
    o = po.Oracle(LX, LY, bc=2, **ZERO_PHYS)
    nb = o.neighbours()
    assert o.T == (LX + 2) * (LY + 2) and nb.shape == (o.T, 9)
    xy = layer_xy(LX, LY)
    index = {tuple(p): i for i, p in enumerate(xy)}
    assert len(index) == o.T
    for v in range(9):
        assert sorted(nb[:, v]) == list(range(o.T))  # a permutation per direction
        for i, (X, Y) in enumerate(xy):
            want = index[((X + C[v, 0]) % (LX + 2), (Y + C[v, 1]) % (LY + 2))]
            assert nb[i, v] == want
    # interior domain nodes see the periodic table's neighbours
    p = po.Oracle(LX, LY, bc=0, **ZERO_PHYS).neighbours()
    for x in range(1, LX - 1):
        for y in range(1, LY - 1):
            assert np.array_equal(nb[y + LY * x], p[y + LY * x])


@pytest.mark.parametrize("bc", BCS)
def test_walls_conserve_mass_and_phi(bc):
    o = po.Oracle(24, 20, npc=2, bc=bc, **ACTIVE)
    o.configure(**CONFIG0_INIT)
    s0 = o.sums()
    o.step(150)
    s1 = o.sums()
    assert abs(s1["fsum"] - s0["fsum"]) < 1e-13 * s0["fsum"]
    assert abs(s1["phisum"] - s0["phisum"]) < 1e-12 * s0["phisum"]
    ux, uy = vel(o.get("ff"))
    assert np.abs(ux).max() > 1e-6  # the active flow is not trivial


@pytest.mark.parametrize("bc", BCS)
def test_rest_state_is_a_fixed_point(bc):
    LX, LY = 9, 7
    o = po.Oracle(LX, LY, npc=2, bc=bc, **dict(CONFIG0_PHYS, zeta=0.0))
    N = LX * LY
    ff = np.tile(po.equilibrium(0.0, 0.0, 40.0), (N, 1))
    o.set_state(ff, np.zeros(N), np.zeros(N), np.full(N, 0.5))
    o.step(20)
    assert np.allclose(o.get("ff"), ff, rtol=0, atol=1e-13)
    assert np.allclose(o.get("phi"), 0.5, rtol=0, atol=1e-15)


def shear_run(bc, LX, LY, profile, steps, npc=2):
    o = po.Oracle(LX, LY, npc=npc, bc=bc, **dict(ZERO_PHYS, tauNem=1.0, tauIso=1.0))
    N = LX * LY
    y = np.arange(N) % LY
    ff = np.stack([po.equilibrium(u, 0.0, 40.0) for u in profile(y)])
    o.set_state(ff, np.zeros(N), np.zeros(N), np.zeros(N))
    amp = []
    for _ in range(steps):
        o.step(1)
        amp.append(vel(o.get("ff"))[0])
    return np.array(amp)


def decay_rate(amp, lo, hi):
    a = np.abs(amp).max(1)
    return -np.polyfit(np.arange(lo, hi), np.log(a[lo:hi]), 1)[0]


def test_free_slip_channel_keeps_plug_flow_no_slip_brakes_it():
    U = 1e-4
    free = shear_run(1, 4, 16, lambda y: U + 0 * y, 100)
    assert np.abs(free[-1] - U).max() < 1e-11 * U
    stuck = shear_run(2, 4, 16, lambda y: U + 0 * y, 100)
    mom = stuck.sum(1)
    assert np.all(np.diff(mom) < 0) and mom[-1] < 0.5 * mom[0]


def test_shear_modes_decay_as_with_walls_half_a_node_out():
    # kinematic viscosity from a periodic shear wave (k = 2 pi / LY), then the
    # lowest mode between walls at y = -1/2 and LY - 1/2 (k = pi / LY): sin
    # for no slip, cos for free slip; walls one node off give ratios of
    # (32/33)^2 = 0.94 or (32/31)^2 = 1.07
    LY, U = 32, 1e-4
    per = shear_run(0, 4, LY, lambda y: U * np.sin(2 * np.pi * (y + .5) / LY), 400)
    nu = decay_rate(per, 100, 400) / (2 * np.pi / LY) ** 2
    assert abs(nu - 1 / 3) < 0.01  # tau/3 for this predictor-corrector scheme
    k2 = (np.pi / LY) ** 2
    ns = shear_run(2, 4, LY, lambda y: U * np.sin(np.pi * (y + .5) / LY), 400)
    fs = shear_run(1, 4, LY, lambda y: U * np.cos(np.pi * (y + .5) / LY), 400)
    assert abs(decay_rate(ns, 100, 400) / (nu * k2) - 1) < 0.02
    assert abs(decay_rate(fs, 100, 400) / (nu * k2) - 1) < 0.02


@pytest.mark.parametrize("bc", [1, 2, 3, 4])
def test_velocity_ghosts_continue_a_linear_profile(bc):
    # with Q = 0 the predictor's Q increment is D_yx = xi (dx uy + dy ux)
    # (lyotropic.cpp:286-295); the velocity ghosts extrapolate linearly
    # (fields.hpp:738-768), so a linear shear ux = g (y - c) gives the same
    # dy ux = g in the rows next to the walls as inside
    LX, LY, g, xi = 6, 10, 1e-4, 0.7
    o = po.Oracle(LX, LY, npc=1, bc=bc, **dict(ZERO_PHYS, xi=xi))
    N = LX * LY
    y = np.arange(N) % LY
    ff = np.stack([po.equilibrium(g * (yy - 4.5), 0.0, 40.0) for yy in y])
    o.set_state(ff, np.zeros(N), np.zeros(N), np.zeros(N))
    o.boundary(0)
    o.update_quantities()
    o.boundary(1)
    o.update_fields(True)
    qn = o.get("QNyx")
    assert np.allclose(qn, 0.5 * xi * g, rtol=1e-9, atol=0)


# symmetry maps of the state: node permutation, direction permutation, signs
MIRROR_X = ([0, 2, 1, 3, 4, 7, 8, 5, 6], 1, -1)   # (cx, cy) -> (-cx, cy); Qxx, Qyx signs
MIRROR_Y = ([0, 1, 2, 4, 3, 8, 7, 6, 5], 1, -1)   # (cx, cy) -> (cx, -cy)
TRANSPOSE = ([0, 3, 4, 1, 2, 5, 6, 8, 7], -1, 1)  # (cx, cy) -> (cy, cx)


def transform(kind, LX, LY, ff, qxx, qyx, phi):
    x, y = np.meshgrid(np.arange(LX), np.arange(LY), indexing="ij")
    x, y = x.reshape(-1), y.reshape(-1)
    if kind == "mx":
        perm, sxx, syx = MIRROR_X
        dst = y + LY * (LX - 1 - x)
    elif kind == "my":
        perm, sxx, syx = MIRROR_Y
        dst = (LY - 1 - y) + LY * x
    elif kind == "t":
        perm, sxx, syx = TRANSPOSE
        dst = x + LX * y
    else:  # translation by 3 along x
        perm, sxx, syx = list(range(9)), 1, 1
        dst = y + LY * ((x + 3) % LX)
    f2 = np.empty_like(ff)
    f2[dst[:, None], np.array(perm)[None, :]] = ff
    out = [f2]
    for a, s in ((qxx, sxx), (qyx, syx), (phi, 1)):
        b = np.empty_like(a)
        b[dst] = s * a
        out.append(b)
    return out


def state(o):
    return [o.get(n) for n in ("ff", "QQxx", "QQyx", "phi")]


@pytest.mark.parametrize("bc,kind", [(3, "mx"), (3, "my"), (3, "t"), (4, "mx"), (4, "my"),
                                     (4, "t"), (1, "mx"), (1, "my"), (2, "mx"), (2, "my")])
def test_walls_commute_with_mirrors_and_transpose(bc, kind):
    L = 12
    a = po.Oracle(L, L, npc=2, bc=bc, **ACTIVE)
    a.configure(**CONFIG0_INIT)
    s0 = state(a)
    b = po.Oracle(L, L, npc=2, bc=bc, **ACTIVE)
    b.set_state(*transform(kind, L, L, *s0))
    a.step(25)
    b.step(25)
    want = transform(kind, L, L, *state(a))
    got = state(b)
    for w, g in zip(want, got):
        assert np.allclose(g, w, rtol=0, atol=1e-11 * max(1.0, np.abs(w).max()))
    # the flow is not symmetric by itself: the check has something to check
    assert np.abs(want[1] - state(a)[1]).max() > 1e-3


@pytest.mark.parametrize("bc", [1, 2])
def test_channel_is_translation_invariant_along_x_bitwise(bc):
    LX, LY = 10, 8
    a = po.Oracle(LX, LY, npc=2, bc=bc, **ACTIVE)
    a.configure(**CONFIG0_INIT)
    b = po.Oracle(LX, LY, npc=2, bc=bc, **ACTIVE)
    b.set_state(*transform("shift", LX, LY, *state(a)))
    a.step(20)
    b.step(20)
    for w, g in zip(transform("shift", LX, LY, *state(a)), state(b)):
        assert np.array_equal(w, g)


def test_wall_initial_configurations():
    # lyotropic.cpp:96-116: with bc 4 the box and the two cutting discs sit
    # at the wall x = 0
    LX, LY, level, radius = 20, 16, 6.0, 3.0
    x, y = np.meshgrid(np.arange(LX), np.arange(LY), indexing="ij")
    x, y = x.reshape(-1), y.reshape(-1)
    o = po.Oracle(LX, LY, bc=4, **dict(ZERO_PHYS, conc=1.0))
    o.configure("boxatwall", seed=1, level=level, noise=0.0)
    assert np.array_equal(o.get("phi") > 0, (x < level) & (np.abs(LY // 2 - y) < level / 2))
    o.configure("cuttingatwall", seed=1, radius=radius, noise=0.0)
    off = int(radius * 1.5)
    dy1 = np.abs(y - (LY // 2 + off)) % LY
    dy2 = np.abs(y - (LY // 2 - off)) % LY
    d1 = np.minimum(dy1, LY - dy1)
    d2 = np.minimum(dy2, LY - dy2)
    want = (x ** 2 + d1 ** 2 <= radius ** 2) | (x ** 2 + d2 ** 2 <= radius ** 2)
    assert np.array_equal(o.get("phi") > 0, want)
    # any other bc keeps the periodic branch
    p = po.Oracle(LX, LY, bc=2, **dict(ZERO_PHYS, conc=1.0))
    p.configure("boxatwall", seed=1, level=level, noise=0.0)
    assert np.array_equal(p.get("phi") > 0, (y < level) & (np.abs(LX // 2 - x) < level))


def test_bad_bc_is_rejected():
    with pytest.raises(ValueError):
        po.Oracle(8, 8, bc=5, **ZERO_PHYS)
