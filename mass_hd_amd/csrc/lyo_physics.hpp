// lyo_physics.hpp -- per-node arithmetic of the lyotropic step, for gfx950.
//
// Two instantiations share this code:
//   EXACT=true,  T=double : the reference's FP64 arithmetic in the reference's
//                           expression order (true divisions, +0*stencil terms,
//                           sequential sums).  Built with -ffp-contract=off so
//                           every operation rounds exactly as the reference's
//                           x86-64 build does.
//   EXACT=false, T=float  : populations stored as deviations f' = f - w_i*rho
//                           from the rest equilibrium, reciprocals instead of
//                           divisions, FMA contraction allowed, and the same
//                           quantities regrouped for fewer operations: raw
//                           central differences with the one-half factors in
//                           per-run constants, y neighbours as sum and
//                           difference, and the collision as even and odd
//                           parts of each opposite population pair (equal to
//                           the reference's expressions in exact arithmetic).
//
// Reference: src/models/lyotropic.cpp:168-341, src/lb.hpp:15-41,73-90,127-136.
#pragma once

#include <hip/hip_runtime.h>

namespace lyo {

template <class T>
struct Params {
  T GammaQ, GammaP, xi, tauNem, tauIso, rho, friction;
  T AA, CC, LL, KK, zeta, zetaI, conc;
  // Per-run constants of the FP32 arithmetic, set by finish_params (the FP64
  // path does not read them): stress coefficients for raw differences
  // (a(+1) - a(-1), twice the central derivative) KK/8, KK/4, LL/4, LL/2
  // (exact power-of-two scalings); zetaI*conc; tauNem - tauIso; friction/3,
  // friction/12; GammaQ/2, friction/6, friction/24 (the halved increments)
  T KK8, KK4, LL4, LL2, zIc, dtau, fr3, fr12, GammaQh, fr3h, fr12h;
};

template <class T>
inline void finish_params(Params<T>& P) {
  P.KK8 = P.KK * T(.125);
  P.KK4 = P.KK * T(.25);
  P.LL4 = P.LL * T(.25);
  P.LL2 = P.LL * T(.5);
  P.zIc = P.zetaI * P.conc;
  P.dtau = P.tauNem - P.tauIso;
  P.fr3 = P.friction * T(1. / 3.);
  P.fr12 = P.friction * T(1. / 12.);
  P.GammaQh = P.GammaQ * T(.5);
  P.fr3h = P.fr3 * T(.5);
  P.fr12h = P.fr12 * T(.5);
}

// Rest-equilibrium weights fe(u=0, n=1) of src/lb.hpp:24-41: these define the
// shift of the FP32 populations (they are not the forcing weights w[]).
__device__ __host__ constexpr double rest_weight(int v) {
  return v == 0 ? 4. / 9. : (v < 5 ? 1. / 9. : 1. / 36.);
}

// ---------------------------------------------------------- stencils
// src/lb.hpp:73-90: derivX/derivY = .5*(a[d1]-a[d2]) + 0*stencil
template <class T, bool EXACT>
__device__ __forceinline__ T dcen(T ap, T am) {
  if constexpr (EXACT) {
    return T(.5) * (ap - am) + T(0);
  } else {
    return T(.5) * (ap - am);
  }
}
// src/lb.hpp:127-136: a[d1]+a[d2]+a[d3]+a[d4] - 4*a[d0] + 0*stencil, with
// d1 = x+1, d2 = x-1, d3 = y+1, d4 = y-1.
template <class T, bool EXACT>
__device__ __forceinline__ T lap5(T c, T xp, T xm, T yp, T ym) {
  if constexpr (EXACT) {
    return xp + xm + yp + ym - T(4) * c + T(0);
  } else {
    return (xp + xm) + (yp + ym) - T(4) * c;
  }
}

// The y+1 and y-1 values of a quantity at a node, as the node arithmetic
// reads them.  FP64 (EXACT) keeps both: the reference's sums run left to
// right over them.  FP32 keeps their sum and difference, all its stencils use,
// formed where the values are taken (from the adjacent lanes of a wave), so
// that the lane shift folds into the add or subtract that consumes it.
template <class T, bool EXACT>
struct YNb;
template <class T>
struct YNb<T, true> {
  T p, m;
};
template <class T>
struct YNb<T, false> {
  T s, d;  // (y+1) + (y-1), (y+1) - (y-1)
};
template <class T, bool EXACT>
__device__ __forceinline__ YNb<T, EXACT> ynb(T yp, T ym) {
  if constexpr (EXACT) {
    return {yp, ym};
  } else {
    return {yp + ym, yp - ym};
  }
}
template <class T, bool EXACT>
__device__ __forceinline__ T lap5(T c, T xp, T xm, const YNb<T, EXACT>& y) {
  if constexpr (EXACT) {
    return lap5<T, true>(c, xp, xm, y.p, y.m);
  } else {
    return (xp + xm) + y.s - T(4) * c;
  }
}

// a / b rounded to nearest from y = RN(1/b), for finite a and finite b > 0
// (Markstein's theorem: with y the correctly rounded reciprocal, q0 =
// RN(a y) is within one ulp of a/b, the residual r = a - q0 b is exact in one
// FMA, and RN(q0 + r y) = RN(a/b)).  The sign of a zero quotient is the
// dividend's (b > 0), which q0 + r y would lose for a = -0.  One IEEE
// division per divisor instead of one per quotient, the same bits.
__device__ __forceinline__ double quot_rn(double a, double b, double y) {
  const double q0 = a * y;
  const double r = __builtin_fma(-q0, b, a);
  return __builtin_copysign(__builtin_fma(r, y, q0), a);
}

template <class T>
__device__ __forceinline__ T fast_rcp(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    return T(1) / x;
  }
}

// Values of one node that its neighbours' UpdateFields reads
// (ux, uy, ux_phi, uy_phi, MU, sigma** of lyotropic.cpp:224-239).
template <class T>
struct NbrVals {
  T ux, uy, uxp, uyp, mu, sxx, syy, sxy, syx;
};
// Values UpdateFields reads at the node itself (lyotropic.cpp:255-266).
// FP32: the Q derivatives are kept as raw differences (twice the central
// derivative); UpdateFields applies the factor one half once per increment.
template <class T>
struct CenterVals {
  T nn, dn, vx, vy, Hxx, Hyx, dxQxx, dyQxx, dxQyx, dyQyx;
};
// Diagnostic stress parts written into frames (lyotropic.cpp:240-246).
template <class T>
struct DiagVals {
  T bulk, el_xx, el_yx, pf_xx, pf_yx, act_xx, act_yx;
};

// UpdateQuantitiesAtNode (src/models/lyotropic.cpp:168-247) for one node.
// f: populations at the node (EXACT: f; else f').  q*/p*: 5-point stencil
// values (centre, x+1, x-1, y neighbours) of Qxx, Qyx, phi.
template <class T, bool EXACT, bool DIAG = false>
__device__ __forceinline__ void update_quantities(
    const T (&f)[9], T qxx_c, T qxx_xp, T qxx_xm, const YNb<T, EXACT>& qxx_y, T qyx_c,
    T qyx_xp, T qyx_xm, const YNb<T, EXACT>& qyx_y, T p_c, T p_xp, T p_xm,
    const YNb<T, EXACT>& p_y, const Params<T>& P, NbrVals<T>& o, CenterVals<T>& c,
    DiagVals<T>* dg = nullptr) {
  const T Qxx = qxx_c, Qyx = qyx_c, p = p_c;
  T nn, dn, vx, vy;
  if constexpr (EXACT) {
    nn = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    vx = (f[1] - f[2] + f[5] - f[6] - f[7] + f[8]) / nn;
    vy = (f[3] - f[4] + f[5] - f[6] + f[7] - f[8]) / nn;
    dn = T(0);
  } else {
    // n = rho + sum f' (the rest weights sum to one); momentum is unchanged
    // by the shift because opposite directions carry equal weights.
    dn = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) + f[8];
    nn = P.rho + dn;
    const T inv = fast_rcp(nn);
    const T d56 = f[5] - f[6], d87 = f[8] - f[7];
    vx = ((f[1] - f[2]) + d56 + d87) * inv;
    vy = ((f[3] - f[4]) + d56 - d87) * inv;
  }

  const T del2Qxx = lap5<T, EXACT>(qxx_c, qxx_xp, qxx_xm, qxx_y);
  const T del2Qyx = lap5<T, EXACT>(qyx_c, qyx_xp, qyx_xm, qyx_y);
  const T del2p = lap5<T, EXACT>(p_c, p_xp, p_xm, p_y);
  const T term = T(0) - Qxx * Qxx - Qyx * Qyx;
  const T Hxx = P.CC * term * Qxx + P.LL * del2Qxx;
  const T Hyx = P.CC * term * Qyx + P.LL * del2Qyx;
  T mu, sigmaB;
  if constexpr (EXACT) {
    mu = P.AA * p * (T(1) - p) * (T(1) - T(2) * p) + T(0) * P.CC * term - P.KK * del2p;
    sigmaB = T(.5) * P.AA * p * p * (T(1) - p) * (T(1) - p) + T(.5) * P.CC * term * term -
             mu * p;
  } else {
    // with g = p (1 - p): mu = AA g (1 - 2p) - KK lap(phi),
    // sigma_bulk = (AA g^2 + CC term^2)/2 - mu p
    const T AAg = P.AA * (p * (T(1) - p));
    mu = AAg * (T(1) - T(2) * p) - P.KK * del2p;
    sigmaB = T(.5) * (AAg * (p * (T(1) - p)) + (P.CC * term) * term) - mu * p;
  }

  // derivatives and the gradient (Ericksen) stresses; FP32 works on raw
  // differences r = 2 d with the coefficients scaled by 1/4 (KK/2 -> KK/8,
  // KK -> KK/4, LL -> LL/4, 2 LL -> LL/2), the same values in exact arithmetic
  T dxQxx, dyQxx, dxQyx, dyQyx, grad_f, grad_s, pf_xx, pf_yx;
  if constexpr (EXACT) {
    dxQxx = dcen<T, EXACT>(qxx_xp, qxx_xm);
    dyQxx = dcen<T, EXACT>(qxx_y.p, qxx_y.m);
    dxQyx = dcen<T, EXACT>(qyx_xp, qyx_xm);
    dyQyx = dcen<T, EXACT>(qyx_y.p, qyx_y.m);
    const T dxPhi = dcen<T, EXACT>(p_xp, p_xm);
    const T dyPhi = dcen<T, EXACT>(p_y.p, p_y.m);
    pf_xx = T(.5) * P.KK * (dyPhi * dyPhi - dxPhi * dxPhi);
    pf_yx = -P.KK * dxPhi * dyPhi;
    // the reference's expressions as written (lyotropic.cpp:212-219)
    const T sigmaF = T(2) * P.xi * ((Qxx * Qxx - T(1)) * Hxx + Qxx * Qyx * Hyx) -
                     P.zeta * Qxx * (T(1) - p) +
                     T(.5) * P.KK * (dyPhi * dyPhi - dxPhi * dxPhi) +
                     P.LL * (dyQxx * dyQxx + dyQyx * dyQyx - dxQxx * dxQxx - dxQyx * dxQyx);
    const T sigmaS = T(2) * P.xi * (Qyx * Qxx * Hxx + (Qyx * Qyx - T(1)) * Hyx) -
                     P.zeta * Qyx * (T(1) - p) - P.KK * dxPhi * dyPhi -
                     T(2) * P.LL * (dxQxx * dyQxx + dxQyx * dyQyx);
    grad_f = sigmaF;
    grad_s = sigmaS;
  } else {
    dxQxx = qxx_xp - qxx_xm;
    dyQxx = qxx_y.d;
    dxQyx = qyx_xp - qyx_xm;
    dyQyx = qyx_y.d;
    const T rxP = p_xp - p_xm, ryP = p_y.d;
    pf_xx = P.KK8 * (ryP * ryP - rxP * rxP);
    pf_yx = -P.KK4 * rxP * ryP;
    grad_f = T(2) * P.xi * ((Qxx * Qxx - T(1)) * Hxx + Qxx * Qyx * Hyx) -
             P.zeta * Qxx * (T(1) - p) + pf_xx +
             P.LL4 * (dyQxx * dyQxx + dyQyx * dyQyx - dxQxx * dxQxx - dxQyx * dxQyx);
    grad_s = T(2) * P.xi * (Qyx * Qxx * Hxx + (Qyx * Qyx - T(1)) * Hyx) -
             P.zeta * Qyx * (T(1) - p) + pf_yx -
             P.LL2 * (dxQxx * dyQxx + dxQyx * dyQyx);
  }
  const T sigmaF = grad_f, sigmaS = grad_s;
  // antisymmetric stress sigmaA = 2 (Qxx Hyx - Qyx Hxx)
  const T rotA = Qxx * Hyx - Qyx * Hxx;

  o.ux = vx;
  o.uy = vy;
  o.uxp = vx * p;
  o.uyp = vy * p;
  o.mu = mu;
  // isotropic part: sigma_bulk + zetaI (conc - phi)
  T iso;
  if constexpr (EXACT) {
    o.sxx = sigmaF + sigmaB + P.zetaI * (P.conc - p);
    o.syy = -sigmaF + sigmaB + P.zetaI * (P.conc - p);
    iso = sigmaB + P.zetaI * (P.conc - p);
  } else {
    iso = sigmaB + (P.zIc - P.zetaI * p);
    o.sxx = iso + sigmaF;
    o.syy = iso - sigmaF;
  }
  if constexpr (EXACT) {
    const T sigmaA = T(2) * rotA;
    o.sxy = sigmaS + sigmaA;
    o.syx = sigmaS - sigmaA;
  } else {
    o.sxy = sigmaS + T(2) * rotA;
    o.syx = sigmaS - T(2) * rotA;
  }

  c.nn = nn;
  c.dn = dn;
  c.vx = vx;
  c.vy = vy;
  c.Hxx = Hxx;
  c.Hyx = Hyx;
  c.dxQxx = dxQxx;
  c.dyQxx = dyQxx;
  c.dxQyx = dxQyx;
  c.dyQyx = dyQyx;

  if constexpr (DIAG) {
    dg->bulk = iso;
    dg->el_xx = sigmaF - pf_xx + P.zeta * Qxx * (T(1) - p);
    dg->el_yx = sigmaS - pf_yx + P.zeta * Qxx * (T(1) - p);
    dg->pf_xx = pf_xx;
    dg->pf_yx = pf_yx;
    dg->act_xx = -P.zeta * Qxx * (T(1) - p);
    dg->act_yx = -P.zeta * Qyx * (T(1) - p);
  }
}

// Increments of UpdateFieldsAtNode (src/models/lyotropic.cpp:268-308): the
// Beris-Edwards D, the Cahn-Hilliard Dp, the force and the collision operator
// Omega_v = (fe_v - f_v)/tau + w_v (F . c_v).
// Stencil inputs: x+1 / x-1 values of ux, uy, ux_phi, sigmaXX, sigmaYX, MU and
// the y neighbours of ux, uy, uy_phi, sigmaXY, sigmaYY, MU.
template <class T, bool EXACT>
struct UFStencil {
  T mu_c;
  T ux_xp, ux_xm, uy_xp, uy_xm, uxp_xp, uxp_xm, sxx_xp, sxx_xm, syx_xp, syx_xm, mu_xp, mu_xm;
  YNb<T, EXACT> ux_y, uy_y, uyp_y, sxy_y, syy_y, mu_y;
};

// POST (FP32 only): om receives the post-collision populations f + Omega
// instead of Omega.  HALF (FP32 only): om, Dxx and Dyx receive Omega/2, D/2
// (every constant of their expressions halved, so the results are exactly
// half of the full ones); Dp stays whole.
template <class T, bool EXACT, bool POST = false, bool HALF = false>
__device__ __forceinline__ void update_fields(const CenterVals<T>& c, T Qxx, T Qyx,
                                              T p, const T (&f)[9],
                                              const UFStencil<T, EXACT>& s,
                                              const Params<T>& P, T& Dxx,
                                              T& Dyx, T& Dp, T (&om)[9]) {
  static_assert(!(EXACT && (POST || HALF)), "the FP64 path keeps the reference's forms");
  static_assert(!(POST && HALF), "one output form");
  const T nn = c.nn, vx = c.vx, vy = c.vy;
  const T Hxx = c.Hxx, Hyx = c.Hyx;
  const T dxQxx = c.dxQxx, dyQxx = c.dyQxx, dxQyx = c.dxQyx, dyQyx = c.dyQyx;

  const T del2mu = lap5<T, EXACT>(s.mu_c, s.mu_xp, s.mu_xm, s.mu_y);
  // EXACT: the force; FP32: twice the stress divergence (raw differences)
  T Fx, Fy, gx, gy;
  if constexpr (EXACT) {
    const T dxux = dcen<T, EXACT>(s.ux_xp, s.ux_xm);
    const T dyux = dcen<T, EXACT>(s.ux_y.p, s.ux_y.m);
    const T dxuy = dcen<T, EXACT>(s.uy_xp, s.uy_xm);
    const T dyuy = dcen<T, EXACT>(s.uy_y.p, s.uy_y.m);
    const T dxSxx = dcen<T, EXACT>(s.sxx_xp, s.sxx_xm);
    const T dySxy = dcen<T, EXACT>(s.sxy_y.p, s.sxy_y.m);
    const T dxSyx = dcen<T, EXACT>(s.syx_xp, s.syx_xm);
    const T dySyy = dcen<T, EXACT>(s.syy_y.p, s.syy_y.m);
    const T pFlux = dcen<T, EXACT>(s.uxp_xp, s.uxp_xm) + dcen<T, EXACT>(s.uyp_y.p, s.uyp_y.m);

    const T expansion = dxux + dyuy;
    const T shear = T(.5) * (dxuy + dyux);
    const T vorticity = T(.5) * (dxuy - dyux);
    const T traceQL = Qxx * (dxux - dyuy) + T(2) * Qyx * shear;

    Dxx = P.GammaQ * Hxx - vx * dxQxx - vy * dyQxx - T(2) * vorticity * Qyx +
          P.xi * ((Qxx + T(1)) * (T(2) * dxux - traceQL) + T(2) * Qyx * shear - expansion);
    Dyx = P.GammaQ * Hyx - vx * dxQyx - vy * dyQyx + T(2) * vorticity * Qxx +
          P.xi * (Qyx * (expansion - traceQL) + T(2) * shear);
    // conserve_phi is false for plain Lyotropic: the correction term is 0
    Dp = P.GammaP * del2mu - pFlux - T(0);
    Fx = dxSxx + dySxy - P.friction * vx;
    Fy = dxSyx + dySyy - P.friction * vy;
  } else {
    // raw differences r = 2 d (the Q derivatives of CenterVals are raw too):
    // E2 = 2 expansion, S4 = 4 shear, W4 = 4 vorticity, TQ2 = 2 traceQL; the
    // flow and advection terms are formed doubled and halved once
    const T rxux = s.ux_xp - s.ux_xm, ryux = s.ux_y.d;
    const T rxuy = s.uy_xp - s.uy_xm, ryuy = s.uy_y.d;
    const T E2 = rxux + ryuy;
    const T S4 = rxuy + ryux;
    const T W4 = rxuy - ryux;
    const T TQ2 = Qxx * (rxux - ryuy) + Qyx * S4;
    const T Vxx = P.xi * ((Qxx + T(1)) * (T(2) * rxux - TQ2) + Qyx * S4 - E2) - W4 * Qyx;
    const T Vyx = P.xi * (Qyx * (E2 - TQ2) + S4) + W4 * Qxx;
    const T axx = vx * dxQxx + vy * dyQxx;
    const T ayx = vx * dxQyx + vy * dyQyx;
    if constexpr (HALF) {
      Dxx = P.GammaQh * Hxx + T(.25) * (Vxx - axx);
      Dyx = P.GammaQh * Hyx + T(.25) * (Vyx - ayx);
    } else {
      Dxx = P.GammaQ * Hxx + T(.5) * (Vxx - axx);
      Dyx = P.GammaQ * Hyx + T(.5) * (Vyx - ayx);
    }
    const T rflux = (s.uxp_xp - s.uxp_xm) + s.uyp_y.d;
    Dp = P.GammaP * del2mu - T(.5) * rflux;
    gx = (s.sxx_xp - s.sxx_xm) + s.sxy_y.d;
    gy = (s.syx_xp - s.syx_xm) + s.syy_y.d;
  }

  if constexpr (EXACT) {
    const T tau = P.tauNem * p + P.tauIso * (T(1) - p);
    // src/lb.hpp:24-41 in the reference order
    T fe[9];
    const T v2 = vx * vx + vy * vy;
    fe[1] = nn * (1. / 9. + 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
    fe[2] = nn * (1. / 9. - 1. / 3. * vx + 1. / 3. * v2 - 1. / 2. * vy * vy);
    fe[3] = nn * (1. / 9. + 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
    fe[4] = nn * (1. / 9. - 1. / 3. * vy - 1. / 2. * vx * vx + 1. / 3. * v2);
    fe[5] = nn * (1. / 36. + 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
    fe[6] = nn * (1. / 36. - 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 + 1. / 4. * vx * vy);
    fe[7] = nn * (1. / 36. - 1. / 12. * vx + 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
    fe[8] = nn * (1. / 36. + 1. / 12. * vx - 1. / 12. * vy + 1. / 12. * v2 - 1. / 4. * vx * vy);
    fe[0] = nn - fe[1] - fe[2] - fe[3] - fe[4] - fe[5] - fe[6] - fe[7] - fe[8];
    // w[v]*(Fx*xdir(v) + Fy*ydir(v)) with w = {0, 1/3 x4, 1/12 x4} (lb.hpp:15-16)
    constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
    constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
    constexpr double wf[9] = {0, 1 / 3., 1 / 3., 1 / 3., 1 / 3., 1 / 12., 1 / 12., 1 / 12., 1 / 12.};
    // (fe_v - f_v) / tau, rounded as the division (tau > 0)
    const T itau = T(1) / tau;
#pragma unroll
    for (int v = 0; v < 9; ++v)
      om[v] = quot_rn(fe[v] - f[v], tau, itau) + wf[v] * (Fx * T(cx[v]) + Fy * T(cy[v]));
  } else {
    // Omega_v = g_v - f'_v / tau with g_v = fe'_v / tau + force_v and the
    // shifted equilibrium fe'_v = fe_v - rest_weight(v) rho
    //   = dn rest_weight(v) + nn E_v(u)   (src/lb.hpp:24-41 regrouped):
    //   3 E_{1,2} = cx +- ux, 3 E_{3,4} = cy +- uy, cx = ux^2 - uy^2/2,
    //   cy = uy^2 - ux^2/2; 12 E_{5,6} = s+ +- (ux + uy), 12 E_{7,8} =
    //   s- -+ (ux - uy), s+- = u^2 +- 3 ux uy; E_0 = 4/9 - (2/3) u^2.
    // The force w_v (F . c_v) with F = (stress divergence) - friction u is
    // odd in c_v like the terms linear in u, so each opposite pair shares an
    // even part a and an odd part l: g = a +- l.
    const T tau = P.tauIso + P.dtau * p;
    const T it = fast_rcp(tau);
    const T itx = HALF ? T(.5) * it : it;  // (HALF: every term below halved)
    const T vx2 = vx * vx, vy2 = vy * vy, v2 = vx2 + vy2;
    const T itnn = itx * nn, itdn = itx * c.dn;
    const T n3 = itnn * T(1. / 3.), n12 = itnn * T(1. / 12.);
    const T d9 = itdn * T(1. / 9.), d36 = itdn * T(1. / 36.);
    const T tx = T(3) * vx;
    const T ax = n3 * (vx2 - T(.5) * vy2) + d9;
    const T ay = n3 * (vy2 - T(.5) * vx2) + d9;
    const T ap = n12 * (tx * vy + v2) + d36;
    const T am = n12 * (v2 - tx * vy) + d36;
    const T fr3 = HALF ? P.fr3h : P.fr3, fr12 = HALF ? P.fr12h : P.fr12;
    const T g6 = HALF ? T(1. / 12.) : T(1. / 6.), g24 = HALF ? T(1. / 48.) : T(1. / 24.);
    const T lx = (n3 - fr3) * vx + g6 * gx;
    const T ly = (n3 - fr3) * vy + g6 * gy;
    const T lp = (n12 - fr12) * (vx + vy) + g24 * (gx + gy);
    const T lm = (n12 - fr12) * (vx - vy) + g24 * (gx - gy);
    T g[9];
    g[0] = T(4. / 9.) * itdn - itnn * (T(2. / 3.) * v2);
    g[1] = ax + lx;
    g[2] = ax - lx;
    g[3] = ay + ly;
    g[4] = ay - ly;
    g[5] = ap + lp;
    g[6] = ap - lp;
    g[7] = am - lm;
    g[8] = am + lm;
    if constexpr (POST) {
      const T keep = T(1) - it;
#pragma unroll
      for (int v = 0; v < 9; ++v) om[v] = f[v] * keep + g[v];
    } else {
#pragma unroll
      for (int v = 0; v < 9; ++v) om[v] = g[v] - f[v] * itx;
    }
  }
}

}  // namespace lyo
