"""Pins the oracle's Beris-Edwards / Cahn-Hilliard terms against independent
statements of the same physics (reference src/models/lyotropic.cpp:196-221
for the stresses, :281-299 for the Q and phi equations).

Each check states a term in a form that is not the reference's expanded
scalar expression:

* the Q equation against the matrix form of the co-rotational derivative,
  S(W, Q) = (xi E + W)(Qt + I/2) + (Qt + I/2)(xi E - W) - 2 xi (Qt + I/2)
  tr(Qt grad u) with Qt = Q/2 (the reference stores Qxx = S cos 2theta, i.e.
  twice the usual Q = S(nn - I/2)) and its trace projected out;
* advection as translation of the pattern, co-rotation as rigid rotation of
  the director at the local vorticity;
* the stress against the tensor form of the Beris-Edwards stress
  (isotropic f - mu phi, Ericksen stresses -d f/d(grad psi) grad psi, the xi
  coupling -xi(H P + P H) + 2 xi P (Qt:Ht) with P = Qt + I/2, and the
  antisymmetric Qt Ht - Ht Qt), where f is the free energy whose derivatives
  give the reference's H and mu;
* thermodynamic consistency: in the continuum the passive stress obeys
  div sigma = -phi grad mu - Ht:grad Qt (plus the curl of the antisymmetric
  part), and the xi terms of stress and Q equation are Onsager reciprocal,
  sigma_xi : E = -Ht : S_xi.

A sign or factor error in advection, co-rotation, flow alignment, the
capillary (KK) or elastic (LL) Ericksen stresses or the bulk (CC) pressure
fails at least one check.  The discrete derivatives (central differences and
the 5-point Laplacian) are pinned on polynomials in test_oracle_pins.py.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

C = np.array([(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1), (-1, 1), (1, -1)])
ZERO = dict(GammaQ=0.0, GammaP=0.0, xi=0.0, tauNem=1.0, tauIso=1.0, rho=40.0,
            friction=0.0, AA=0.0, CC=0.0, LL=0.0, KK=0.0, zeta=0.0, zetaI=0.0, conc=0.0)


def phys(**kw):
    p = dict(ZERO)
    p.update(kw)
    return p


# ------------------------------------------------------------ grid helpers
# arrays are shaped (LX, LY); flattening in C order gives k = y + LY*x

def dx(a):
    return 0.5 * (np.roll(a, -1, 0) - np.roll(a, 1, 0))


def dy(a):
    return 0.5 * (np.roll(a, -1, 1) - np.roll(a, 1, 1))


def lap(a):
    return np.roll(a, -1, 0) + np.roll(a, 1, 0) + np.roll(a, -1, 1) + np.roll(a, 1, 1) - 4 * a


def smooth_field(LX, LY, rng, amp, modes=((1, 0), (0, 1), (1, 1), (2, -1))):
    """Sum of low Fourier modes with random amplitudes and phases."""
    x, y = np.meshgrid(np.arange(LX), np.arange(LY), indexing="ij")
    out = np.zeros((LX, LY))
    for mx, my in modes:
        a = amp * rng.uniform(0.3, 1.0)
        out += a * np.sin(2 * math.pi * (mx * x / LX + my * y / LY) + rng.uniform(0, 2 * math.pi))
    return out


def flow_state(ux, uy, rho=40.0):
    """Populations at equilibrium with the given velocity field."""
    return np.stack([po.equilibrium(a, b, rho) for a, b in zip(ux.ravel(), uy.ravel())])


def oracle_with(LX, LY, P, ff, qxx, qyx, phi):
    o = po.Oracle(LX, LY, **P)
    o.set_state(ff.reshape(-1), qxx.ravel(), qyx.ravel(), phi.ravel())
    return o


def field(o, name, LX, LY):
    return o.get(name).reshape(LX, LY)


def q_rate(o, qxx, qyx, LX, LY):
    """D of the predictor (QN = Q + D/2) as a per-node (xx, yx) pair."""
    o.update_quantities()
    o.update_fields(True)
    return (2 * (field(o, "QNxx", LX, LY) - qxx), 2 * (field(o, "QNyx", LX, LY) - qyx))


def mat(xx, yx):
    """Traceless symmetric 2x2 tensors from their (xx, yx) components."""
    m = np.empty(xx.shape + (2, 2))
    m[..., 0, 0], m[..., 0, 1], m[..., 1, 0], m[..., 1, 1] = xx, yx, yx, -xx
    return m


def contract(a, b):
    return np.einsum("...ij,...ij->...", a, b)


def mm(a, b):
    return np.einsum("...ij,...jk->...ik", a, b)


EYE = np.eye(2)


def co_rotational(grad_u, qt, xi):
    """S(W, Qt) of the Beris-Edwards equation for the usual Qt, with
    (grad u)_ab = d_b u_a, E and W its symmetric and antisymmetric parts, and
    the trace removed (the lattice flow is weakly compressible)."""
    E = 0.5 * (grad_u + np.swapaxes(grad_u, -1, -2))
    W = 0.5 * (grad_u - np.swapaxes(grad_u, -1, -2))
    P = qt + 0.5 * EYE
    S = (mm(xi * E + W, P) + mm(P, xi * E - W)
         - 2 * xi * P * contract(qt, grad_u)[..., None, None])
    return S - 0.5 * np.trace(S, axis1=-2, axis2=-1)[..., None, None] * EYE


def discrete_grad_u(ux, uy):
    g = np.empty(ux.shape + (2, 2))
    g[..., 0, 0], g[..., 0, 1] = dx(ux), dy(ux)
    g[..., 1, 0], g[..., 1, 1] = dx(uy), dy(uy)
    return g


def molecular_field(qxx, qyx, LL, CC, lap_xx, lap_yx):
    """H = -dF/dQ for F = (LL/2)(|grad Qxx|^2 + |grad Qyx|^2) (per independent
    component; the tensor-normalised Frank energy in the reference's Q) plus the
    quartic bulk term: H = LL lap Q - CC |Q|^2 Q."""
    q2 = qxx * qxx + qyx * qyx
    return LL * lap_xx - CC * q2 * qxx, LL * lap_yx - CC * q2 * qyx


# -------------------------------------------------------- the Q equation

def test_q_advection_translates_the_pattern():
    """Uniform flow u0, no relaxation, no alignment: the predictor moves Q and
    phi with the flow, dQ/dt = -u0 . grad Q (lyotropic.cpp:292,295)."""
    LX, LY = 24, 20
    rng = np.random.default_rng(11)
    qxx, qyx = smooth_field(LX, LY, rng, 0.3), smooth_field(LX, LY, rng, 0.3)
    phi = 0.5 + smooth_field(LX, LY, rng, 0.05)
    u0 = (0.021, -0.013)
    ff = flow_state(np.full((LX, LY), u0[0]), np.full((LX, LY), u0[1]))
    o = oracle_with(LX, LY, phys(LL=0.3), ff, qxx, qyx, phi)
    Dxx, Dyx = q_rate(o, qxx, qyx, LX, LY)
    for q, D in ((qxx, Dxx), (qyx, Dyx)):
        exp = -(u0[0] * dx(q) + u0[1] * dy(q))
        assert np.abs(D).max() > 1e-3
        assert np.allclose(D, exp, rtol=0, atol=1e-15)
    # phi rides along too (GammaP = 0): phn = phi - u0.grad(phi)/2
    exp = phi - 0.5 * (u0[0] * dx(phi) + u0[1] * dy(phi))
    assert np.allclose(field(o, "phn", LX, LY), exp, rtol=0, atol=1e-15)


def test_director_co_rotates_with_the_vorticity():
    """Uniform Q in a divergence-free flow with xi = 0 and no relaxation: the
    director turns rigidly at the local vorticity w = (dx uy - dy ux)/2,
    keeping |Q| (lyotropic.cpp:287,292,295).  With Qxx + i Qyx = s e^{2i
    theta} that is dQxx = -2 w Qyx, dQyx = +2 w Qxx."""
    LX, LY = 20, 24
    rng = np.random.default_rng(5)
    psi = smooth_field(LX, LY, rng, 0.4)
    ux, uy = dy(psi), -dx(psi)          # discrete stream function: div u = 0
    ff = flow_state(ux, uy)
    q0 = (0.27, -0.34)
    qxx, qyx = np.full((LX, LY), q0[0]), np.full((LX, LY), q0[1])
    o = oracle_with(LX, LY, phys(), ff, qxx, qyx, np.full((LX, LY), 0.4))
    Dxx, Dyx = q_rate(o, qxx, qyx, LX, LY)
    ux_o, uy_o = field(o, "ux", LX, LY), field(o, "uy", LX, LY)
    w = 0.5 * (dx(uy_o) - dy(ux_o))
    assert np.abs(w).max() > 1e-3
    q2 = q0[0] ** 2 + q0[1] ** 2
    # a rotation: |Q| unchanged to first order
    assert np.allclose(q0[0] * Dxx + q0[1] * Dyx, 0, atol=1e-16)
    # turning rate of theta (Q ~ e^{2 i theta}) equals the fluid's rotation
    assert np.allclose((q0[0] * Dyx - q0[1] * Dxx) / (2 * q2), w, rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", [1, 2])
def test_q_equation_matches_the_tensor_form(seed):
    """General smooth Q and compressible u, all Q-equation terms on:
    D = GammaQ H - u.grad Q + 2 S(W, Q/2) node by node, with S the matrix
    form above (lyotropic.cpp:281-297)."""
    LX, LY = 22, 18
    rng = np.random.default_rng(seed)
    qxx, qyx = smooth_field(LX, LY, rng, 0.5), smooth_field(LX, LY, rng, 0.5)
    ux, uy = smooth_field(LX, LY, rng, 0.04), smooth_field(LX, LY, rng, 0.04)
    G, xi, LL, CC = 0.3, 0.7, 0.06, 0.4
    o = oracle_with(LX, LY, phys(GammaQ=G, xi=xi, LL=LL, CC=CC), flow_state(ux, uy),
                    qxx, qyx, np.full((LX, LY), 0.5))
    Dxx, Dyx = q_rate(o, qxx, qyx, LX, LY)
    ux_o, uy_o = field(o, "ux", LX, LY), field(o, "uy", LX, LY)
    Hxx, Hyx = molecular_field(qxx, qyx, LL, CC, lap(qxx), lap(qyx))
    S = co_rotational(discrete_grad_u(ux_o, uy_o), 0.5 * mat(qxx, qyx), xi)
    for q, H, s, D in ((qxx, Hxx, S[..., 0, 0], Dxx), (qyx, Hyx, S[..., 1, 0], Dyx)):
        exp = G * H - ux_o * dx(q) - uy_o * dy(q) + 2 * s
        assert np.abs(2 * s).max() > 1e-2
        assert np.allclose(D, exp, rtol=0, atol=2e-15)


# ------------------------------------------------------------ the stress

def free_energy_density(phi, qxx, qyx, AA, CC):
    """Local part of the free energy whose derivatives the reference uses:
    mu_local = AA p(1-p)(1-2p) = dV/dphi, H_bulk = -CC |Q|^2 Q.  The
    quartic coefficient in the tensor normalisation (Qt:Qt)^2 is fixed by
    requiring -dF/dQt = Ht = 2H (checked numerically in the test)."""
    V = 0.5 * AA * phi ** 2 * (1 - phi) ** 2
    qt2 = 0.5 * (qxx * qxx + qyx * qyx)        # Qt:Qt with Qt = Q/2
    return V + 2.0 * CC * qt2 ** 2


def tensor_stress(qxx, qyx, phi, grads, P):
    """Beris-Edwards + Cahn-Hilliard stress in the usual normalisation
    (Qt = Q/2, Ht = -dF/dQt = 2H):
        sigma = (f_local - mu phi + zetaI (conc - phi)) I
                - KK (grad phi grad phi - |grad phi|^2 I / 2)          capillary
                - Lt (d_i Qt_kl d_j Qt_kl - |grad Qt|^2 delta_ij / 2)   Ericksen (Frank)
                - xi (Ht P + P Ht) + 2 xi P (Qt : Ht)                  alignment
                + Qt Ht - Ht Qt                                         antisymmetric
                - zeta (1 - phi) Q                                      active
    with P = Qt + I/2 and Lt = 4 LL (Ht = Lt lap Qt equals 2 LL lap Q).  The
    gradient energies enter through their traceless Ericksen parts, which in
    two dimensions equal f_grad I - d f/d(grad psi) grad psi."""
    dphi_x, dphi_y, dq = grads
    mu_loc = P["AA"] * phi * (1 - phi) * (1 - 2 * phi)
    mu = mu_loc - P["KK"] * lap(phi)
    Hxx, Hyx = molecular_field(qxx, qyx, P["LL"], P["CC"], lap(qxx), lap(qyx))
    Q = mat(qxx, qyx)
    Qt, Ht = 0.5 * Q, 2.0 * mat(Hxx, Hyx)
    Pm = Qt + 0.5 * EYE
    iso = free_energy_density(phi, qxx, qyx, P["AA"], P["CC"]) - mu * phi \
        + P["zetaI"] * (P["conc"] - phi)
    gphi = np.stack([dphi_x, dphi_y], -1)
    cap = -P["KK"] * (np.einsum("...i,...j->...ij", gphi, gphi)
                      - 0.5 * (gphi * gphi).sum(-1)[..., None, None] * EYE)
    # dq[a][i]: derivative along i of Qt component a, summed over the tensor
    gQt = 0.5 * np.stack([mat(dq[0][0], dq[1][0]), mat(dq[0][1], dq[1][1])], -3)  # (...,i,k,l)
    GG = np.einsum("...ikl,...jkl->...ij", gQt, gQt)
    eric = -4 * P["LL"] * (GG - 0.5 * np.trace(GG, axis1=-2, axis2=-1)[..., None, None] * EYE)
    xi = P["xi"]
    align = -xi * (mm(Ht, Pm) + mm(Pm, Ht)) + 2 * xi * Pm * contract(Qt, Ht)[..., None, None]
    anti = mm(Qt, Ht) - mm(Ht, Qt)
    act = -P["zeta"] * (1 - phi)[..., None, None] * Q
    parts = dict(iso=iso[..., None, None] * EYE, cap=cap, eric=eric, align=align, anti=anti,
                 act=act)
    return sum(parts.values()), parts, mu, (Hxx, Hyx)


def test_bulk_free_energy_derivative_gives_the_molecular_field():
    """The quartic free energy used for the isotropic stress is the one whose
    derivative is the Q equation's bulk molecular field, Ht = -dF/dQt = 2H.
    Qxx = Qt_xx - Qt_yy and Qyx = Qt_xy + Qt_yx, so the tensor derivative's
    xx and yx entries are the derivatives in Qxx and Qyx (central finite
    differences here)."""
    rng = np.random.default_rng(3)
    qxx, qyx = rng.uniform(-0.6, 0.6, 5), rng.uniform(-0.6, 0.6, 5)
    CC, h = 0.7, 1e-6
    Hxx, Hyx = molecular_field(qxx, qyx, 0.0, CC, 0, 0)
    f = lambda a, b: free_energy_density(np.zeros_like(a), a, b, 0.0, CC)  # noqa: E731
    dfx = (f(qxx + h, qyx) - f(qxx - h, qyx)) / (2 * h)
    dfy = (f(qxx, qyx + h) - f(qxx, qyx - h)) / (2 * h)
    assert np.allclose(-dfx, 2 * Hxx, rtol=1e-8)
    assert np.allclose(-dfy, 2 * Hyx, rtol=1e-8)


@pytest.mark.parametrize("seed", [4, 9])
def test_stress_matches_the_tensor_form(seed):
    """Every stored stress field of UpdateQuantities (lyotropic.cpp:209-246)
    equals the tensor form node by node, with all couplings on."""
    LX, LY = 20, 22
    rng = np.random.default_rng(seed)
    qxx, qyx = smooth_field(LX, LY, rng, 0.5), smooth_field(LX, LY, rng, 0.5)
    phi = 0.5 + smooth_field(LX, LY, rng, 0.3)
    P = phys(xi=0.7, zeta=0.05, AA=0.2, CC=0.3, LL=0.08, KK=0.1, zetaI=0.03, conc=0.6)
    o = oracle_with(LX, LY, P, flow_state(np.zeros((LX, LY)), np.zeros((LX, LY))),
                    qxx, qyx, phi)
    o.update_quantities()
    grads = (dx(phi), dy(phi), ((dx(qxx), dy(qxx)), (dx(qyx), dy(qyx))))
    sig, parts, mu, (Hxx, Hyx) = tensor_stress(qxx, qyx, phi, grads, P)
    g = lambda n: field(o, n, LX, LY)  # noqa: E731
    assert np.allclose(g("MU"), mu, rtol=0, atol=1e-15)
    assert np.allclose(g("HHxx"), Hxx, rtol=0, atol=1e-15)
    scale = np.abs(sig).max()
    tol = 1e-14 * scale
    assert np.allclose(g("sigmaXX"), sig[..., 0, 0], rtol=0, atol=tol)
    assert np.allclose(g("sigmaYY"), sig[..., 1, 1], rtol=0, atol=tol)
    assert np.allclose(g("sigmaXY"), sig[..., 0, 1], rtol=0, atol=tol)
    assert np.allclose(g("sigmaYX"), sig[..., 1, 0], rtol=0, atol=tol)
    # the frame's split of the stress (lyotropic.cpp:240-246)
    assert np.allclose(g("sigma_bulk"), parts["iso"][..., 0, 0], rtol=0, atol=tol)
    assert np.allclose(g("sigma_phase_field_xx"), parts["cap"][..., 0, 0], rtol=0, atol=tol)
    assert np.allclose(g("sigma_phase_field_yx"), parts["cap"][..., 1, 0], rtol=0, atol=tol)
    assert np.allclose(g("sigma_active_xx"), parts["act"][..., 0, 0], rtol=0, atol=tol)
    assert np.allclose(g("sigma_active_yx"), parts["act"][..., 1, 0], rtol=0, atol=tol)
    el = parts["eric"] + parts["align"]
    assert np.allclose(g("sigma_elastic_xx"), el[..., 0, 0], rtol=0, atol=tol)
    # the reference's elastic yx field adds zeta Qxx (1 - phi) where the
    # active yx stress has Qyx (lyotropic.cpp:243), reproduced as written
    quirk = P["zeta"] * (1 - phi) * (qxx - qyx)
    assert np.allclose(g("sigma_elastic_yx"), el[..., 1, 0] + quirk, rtol=0, atol=tol)
    # every part is far above the tolerance (a sign error in it shows)
    for name in ("cap", "eric", "align", "anti", "act"):
        assert np.abs(parts[name]).max() > 1e6 * tol, name


def spectral(a, LX, LY):
    """Continuum derivatives d/dx, d/dy of a band-limited periodic field."""
    kx = 2 * math.pi * np.fft.fftfreq(LX)[:, None]
    ky = 2 * math.pi * np.fft.fftfreq(LY)[None, :]
    A = np.fft.fft2(a)
    return np.real(np.fft.ifft2(1j * kx * A)), np.real(np.fft.ifft2(1j * ky * A))


def test_passive_force_is_thermodynamic():
    """xi = zeta = zetaI = 0: in the continuum the passive stress satisfies
    div sigma_sym = -phi grad mu - Ht : grad Qt (the Gibbs-Duhem identity of
    the free energy), and the antisymmetric part adds its curl.  The momentum
    each node gains in the predictor collision (F = div sigma with central
    differences, lyotropic.cpp:272-303) must agree to the stencils' O(k^2)
    error on long waves; each coupling (capillary KK, Frank LL, bulk CC,
    double well AA) is large enough that a wrong sign or factor in its stress
    moves F by far more than that."""
    LX = LY = 96
    rng = np.random.default_rng(21)
    modes = ((1, 0), (0, 1), (1, 1), (1, -1))
    qxx = 0.15 + smooth_field(LX, LY, rng, 0.3, modes)
    qyx = smooth_field(LX, LY, rng, 0.3, modes)
    phi = 0.5 + smooth_field(LX, LY, rng, 0.25, modes)
    P = phys(AA=0.05, CC=0.05, LL=4.0, KK=8.0, tauNem=1.3, tauIso=1.3)
    ff0 = flow_state(np.zeros((LX, LY)), np.zeros((LX, LY)))
    o = oracle_with(LX, LY, P, ff0, qxx, qyx, phi)
    o.update_quantities()
    o.update_fields(True)
    dff = o.get("ff") - ff0
    F = [(dff * C[:, a]).sum(1).reshape(LX, LY) for a in (0, 1)]

    # continuum fields (spectrally exact derivatives of the same samples)
    d = lambda a: spectral(a, LX, LY)  # noqa: E731
    def lap_c(a):
        ax, ay = d(a)
        return d(ax)[0] + d(ay)[1]
    mu = P["AA"] * phi * (1 - phi) * (1 - 2 * phi) - P["KK"] * lap_c(phi)
    Hxx, Hyx = molecular_field(qxx, qyx, P["LL"], P["CC"], lap_c(qxx), lap_c(qyx))
    sA = 2 * (qxx * Hyx - qyx * Hxx)                # (Qt Ht - Ht Qt)_xy
    mux, muy = d(mu)
    qxx_x, qxx_y = d(qxx)
    qyx_x, qyx_y = d(qyx)
    sAx, sAy = d(sA)
    # Ht : grad Qt = 2H : grad (Q/2) = H:grad Q = 2 (Hxx dQxx + Hyx dQyx)
    exp_x = -phi * mux - 2 * (Hxx * qxx_x + Hyx * qyx_x) + sAy
    exp_y = -phi * muy - 2 * (Hxx * qxx_y + Hyx * qyx_y) - sAx
    scale = max(np.abs(exp_x).max(), np.abs(exp_y).max())
    err = max(np.abs(F[0] - exp_x).max(), np.abs(F[1] - exp_y).max()) / scale
    assert err < 0.02, err   # 0.4% here: the stencils' O(k^2) error
    # each coupling's share of the force, far above that error
    shares = {
        "phase field": np.abs(phi * mux).max(),
        "bulk CC": np.abs(2 * P["CC"] * (qxx ** 2 + qyx ** 2) * (qxx * qxx_x + qyx * qyx_x)).max(),
        "Frank LL": np.abs(2 * P["LL"] * (lap_c(qxx) * qxx_x + lap_c(qyx) * qyx_x)).max(),
        "antisymmetric": np.abs(sAy).max(),
    }
    for k, v in shares.items():
        assert v > 0.3 * scale, (k, v / scale)


def test_alignment_stress_is_reciprocal_to_flow_alignment():
    """Onsager reciprocity of the xi couplings: at every node the work rate of
    the xi stress on an incompressible strain E equals minus the free-energy
    rate of the xi part of the Q equation, sigma_xi : E = -Ht : S_xi =
    -H : D_xi (Ht = 2H, S_xi = D_xi / 2).  Both sides come from the oracle:
    the xi parts are the differences of runs with and without xi."""
    LX, LY = 18, 20
    rng = np.random.default_rng(8)
    qxx, qyx = smooth_field(LX, LY, rng, 0.5), smooth_field(LX, LY, rng, 0.5)
    phi = np.full((LX, LY), 0.5)
    psi = smooth_field(LX, LY, rng, 0.5)
    ux, uy = dy(psi), -dx(psi)
    base = dict(LL=0.1, CC=0.3)
    rest = flow_state(np.zeros((LX, LY)), np.zeros((LX, LY)))
    flow = flow_state(ux, uy)
    sig = {}
    for xi in (0.0, 0.8):
        o = oracle_with(LX, LY, phys(xi=xi, **base), rest, qxx, qyx, phi)
        o.update_quantities()
        sig[xi] = [field(o, n, LX, LY) for n in ("sigmaXX", "sigmaYY", "sigmaXY", "sigmaYX")]
        H = (field(o, "HHxx", LX, LY), field(o, "HHyx", LX, LY))
    D = {}
    for xi in (0.0, 0.8):
        o = oracle_with(LX, LY, phys(xi=xi, **base), flow, qxx, qyx, phi)
        D[xi] = q_rate(o, qxx, qyx, LX, LY)
        uo = (field(o, "ux", LX, LY), field(o, "uy", LX, LY))
    sxx, syy, sxy, syx = [a - b for a, b in zip(sig[0.8], sig[0.0])]
    Dxx, Dyx = [a - b for a, b in zip(D[0.8], D[0.0])]
    Exx, Eyy = dx(uo[0]), dy(uo[1])
    Exy = 0.5 * (dy(uo[0]) + dx(uo[1]))
    assert np.abs(Exx + Eyy).max() < 1e-15
    work = sxx * Exx + syy * Eyy + (sxy + syx) * Exy
    free = -2 * (H[0] * Dxx + H[1] * Dyx)
    assert np.abs(work).max() > 1e-4
    assert np.allclose(work, free, rtol=0, atol=1e-12 * np.abs(work).max())
