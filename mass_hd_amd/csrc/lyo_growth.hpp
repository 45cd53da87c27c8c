// lyo_growth.hpp -- go-or-grow (src/models/go-or-grow.cpp, with
// src/models/lyotropic-with-division.cpp): node kernels, FP64 in the
// reference's expression order, on the ghost-layer grid of lyo_walls.hpp.
//
// The model adds a phenotype density m (0 <= m <= phi) and the phenotype
// chi = m/phi to Lyotropic: UpdateQuantities takes a crowding penalty B, a
// preferred nematic order Snem and an activity that grows with 1 - chi; after
// every UpdateQuantities and UpdateFields chi is recomputed from m and phi;
// UpdateFields moves m by its own fluxes (chemical-potential drift weighted
// by chi, phenotype diffusion, advection), a switching potential and the
// growth source R = chi*phi*(alpha*division - beta*death) that phi also gets.
// The division/death masks are redrawn on the host (the reference's RNG
// stream) and marked on the device.
#pragma once

#include <hip/hip_runtime.h>

#include "lyo_loops.hpp"
#include "lyo_physics.hpp"

namespace lyo {

// go-or-grow.hpp:10-28, lyotropic-with-division.hpp:19-34 (device copy)
struct GrowthDev {
  double B, Snem, Dchi, Achi, Ochi, phiswitch;
  double alpha, beta, phi_critical;
  int growTogether;
};

// GoOrGrow::UpdateQuantitiesAtNode (go-or-grow.cpp:224-303) for one node:
// the 23 derived values in the order of the loop planes (LoopPlane), from the
// node's populations, the five-point values (centre, x+1, x-1, y+1, y-1) of
// Qxx, Qyx and phi, and chi at the node.
__device__ __forceinline__ void growth_quantities(const double (&fl)[9], double Qxx, double qxx_xp,
                                                  double qxx_xm, double qxx_yp, double qxx_ym,
                                                  double Qyx, double qyx_xp, double qyx_xm,
                                                  double qyx_yp, double qyx_ym, double p,
                                                  double p_xp, double p_xm, double p_yp,
                                                  double p_ym, double chik,
                                                  const Params<double>& P, const GrowthDev& g,
                                                  double (&out)[kLoopPlanes]) {
  using T = double;
  const T nn = fl[0] + fl[1] + fl[2] + fl[3] + fl[4] + fl[5] + fl[6] + fl[7] + fl[8];
  const T vx = (fl[1] - fl[2] + fl[5] - fl[6] - fl[7] + fl[8]) / nn;
  const T vy = (fl[3] - fl[4] + fl[5] - fl[6] + fl[7] - fl[8]) / nn;

  const T del2Qxx = lap5<T, true>(Qxx, qxx_xp, qxx_xm, qxx_yp, qxx_ym);
  const T dxQxx = dcen<T, true>(qxx_xp, qxx_xm);
  const T dyQxx = dcen<T, true>(qxx_yp, qxx_ym);
  const T del2Qyx = lap5<T, true>(Qyx, qyx_xp, qyx_xm, qyx_yp, qyx_ym);
  const T dxQyx = dcen<T, true>(qyx_xp, qyx_xm);
  const T dyQyx = dcen<T, true>(qyx_yp, qyx_ym);
  const T del2p = lap5<T, true>(p, p_xp, p_xm, p_yp, p_ym);
  const T dxPhi = dcen<T, true>(p_xp, p_xm);
  const T dyPhi = dcen<T, true>(p_yp, p_ym);

  const T q2 = Qxx * Qxx + Qyx * Qyx;
  const T term = g.Snem - q2;
  const T Hxx = T(2) * P.CC * term * Qxx + P.LL * del2Qxx;
  const T Hyx = T(2) * P.CC * term * Qyx + P.LL * del2Qyx;
  const T dp_critical = p - g.phi_critical;
  const T f_compress = dp_critical > 0 ? T(.5) * g.B * dp_critical * dp_critical : T(0.);
  const T mu_compress = dp_critical > 0 ? g.B * dp_critical : T(0.);
  const T mu = P.AA * p * (T(1) - p) * (T(1) - T(2) * p) + mu_compress - P.KK * del2p;

  const T sigmaB = T(.5) * P.AA * p * p * (T(1) - p) * (T(1) - p) + f_compress +
                   T(.5) * P.CC * term * term - mu * p;
  const T ap = P.zeta * p * (T(1) - chik);
  const T sigmaF = T(2) * P.xi * ((Qxx * Qxx - T(1)) * Hxx + Qxx * Qyx * Hyx) - ap * Qxx +
                   T(.5) * P.KK * (dyPhi * dyPhi - dxPhi * dxPhi) +
                   P.LL * (dyQxx * dyQxx + dyQyx * dyQyx - dxQxx * dxQxx - dxQyx * dxQyx);
  const T sigmaS = T(2) * P.xi * (Qyx * Qxx * Hxx + (Qyx * Qyx - T(1)) * Hyx) - ap * Qyx -
                   P.KK * dxPhi * dyPhi - T(2) * P.LL * (dxQxx * dyQxx + dxQyx * dyQyx);
  const T sigmaA = T(2) * (Qxx * Hyx - Qyx * Hxx);

  out[LP_N] = nn;
  out[LP_UX] = vx;
  out[LP_UY] = vy;
  out[LP_UXP] = vx * p;
  out[LP_UYP] = vy * p;
  out[LP_HXX] = Hxx;
  out[LP_HYX] = Hyx;
  out[LP_MU] = mu;
  out[LP_DXQXX] = dxQxx;
  out[LP_DXQYX] = dxQyx;
  out[LP_DYQXX] = dyQxx;
  out[LP_DYQYX] = dyQyx;
  out[LP_SXX] = sigmaF + sigmaB + P.zetaI * (P.conc - p);
  out[LP_SYY] = -sigmaF + sigmaB + P.zetaI * (P.conc - p);
  out[LP_SXY] = sigmaS + sigmaA;
  out[LP_SYX] = sigmaS - sigmaA;
  out[LP_BULK] = sigmaB + P.zetaI * (P.conc - p);
  out[LP_ELXX] = sigmaF - T(.5) * P.KK * (dyPhi * dyPhi - dxPhi * dxPhi) + ap * Qxx;
  out[LP_ELYX] = sigmaS + P.KK * dxPhi * dyPhi + ap * Qyx;
  out[LP_PFXX] = T(.5) * P.KK * (dyPhi * dyPhi - dxPhi * dxPhi);
  out[LP_PFYX] = -P.KK * dxPhi * dyPhi;
  out[LP_ACTXX] = -ap * Qxx;
  out[LP_ACTYX] = -ap * Qyx;
}

// GoOrGrow::UpdateQuantitiesAtNode (go-or-grow.cpp:224-303): the 23 derived
// planes of one domain node.
template <class G = LayerGrid>
__global__ void __launch_bounds__(256)
growth_uq_kernel(const double* __restrict__ f, const double* __restrict__ qxx,
                 const double* __restrict__ qyx, const double* __restrict__ phi,
                 const double* __restrict__ chi, Params<double> P, GrowthDev g, int LX, int LY,
                 int NP, double* __restrict__ D) {
  using T = double;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const int k = n.k;
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = f[v * NP + k];
  T out[kLoopPlanes];
  growth_quantities(fl, qxx[k], qxx[n.kxp], qxx[n.kxm], qxx[n.kyp], qxx[n.kym], qyx[k],
                    qyx[n.kxp], qyx[n.kxm], qyx[n.kyp], qyx[n.kym], phi[k], phi[n.kxp],
                    phi[n.kxm], phi[n.kyp], phi[n.kym], chi[k], P, g, out);
#pragma unroll
  for (int i = 0; i < kLoopPlanes; ++i) D[(size_t)i * NP + k] = out[i];
}

// The phenotype density's increment Dm of GoOrGrow::UpdateFields
// (go-or-grow.cpp:324-344) at one node: chi, phi and MU at the node and its
// four neighbours (x+1, x-1, y+1, y-1), m and the velocity component along
// the axis at the neighbours, and the growth source R.
__device__ __forceinline__ double growth_dm(double chi_eff, double chi_px, double chi_mx,
                                            double chi_py, double chi_my, double p, double p_xp,
                                            double p_xm, double p_yp, double p_ym, double m_xp,
                                            double m_xm, double m_yp, double m_ym, double ux_xp,
                                            double ux_xm, double uy_yp, double uy_ym, double mu,
                                            double mu_xp, double mu_xm, double mu_yp,
                                            double mu_ym, double R, const Params<double>& P,
                                            const GrowthDev& g) {
  using T = double;
  const T mFlux = T(.5) * (ux_xp * m_xp - ux_xm * m_xm) + T(.5) * (uy_yp * m_yp - uy_ym * m_ym);
  const T diffusiveMFlux = T(.5) * (chi_px + chi_eff) * (mu_xp - mu) -
                           T(.5) * (chi_eff + chi_mx) * (mu - mu_xm) +
                           T(.5) * (chi_py + chi_eff) * (mu_yp - mu) -
                           T(.5) * (chi_eff + chi_my) * (mu - mu_ym);
  const T phenotypeDiffusion = T(.5) * (p_xp + p) * (chi_px - chi_eff) -
                               T(.5) * (p + p_xm) * (chi_eff - chi_mx) +
                               T(.5) * (p_yp + p) * (chi_py - chi_eff) -
                               T(.5) * (p + p_ym) * (chi_eff - chi_my);
  const T dVdChi = T(2) * g.Achi * chi_eff * (T(1) - chi_eff) * (T(1) - T(2) * chi_eff) +
                   g.Ochi * (p - g.phiswitch);
  const T Sswitch = -p * dVdChi;
  const T mGrowth = g.growTogether ? chi_eff * R : R;
  return P.GammaP * diffusiveMFlux + g.Dchi * phenotypeDiffusion - mFlux + Sswitch + mGrowth;
}

// Inputs and outputs of one go-or-grow UpdateFields loop: Lyotropic's
// (lyo_loops.hpp) plus m, chi, the masks and the phenotype outputs.
struct GrowthUF {
  LoopUF<double> l;
  const double* m;
  const double* chi;
  const double* mn;        // corrector: m + Dm/2 of the predictor
  const unsigned char* division;
  const unsigned char* death;
  double* m_out;
  double* mn_out;          // predictor
  double* chi_out;         // if set: chi = m/phi of the outputs (the phenotype pass)
  const double* countphi;  // conserve_phi: device sum of phi at the last UpdateQuantities
  double totalphi;
  int conserve;
};

// GoOrGrow::UpdateFields (go-or-grow.cpp:322-362) for one domain node; with
// DIV, LyotropicWithDivision::UpdateFields (lyotropic-with-division.cpp:
// 110-126): Lyotropic's update and phi += phi*(alpha*division - beta*death).
template <bool FIRST, class G = LayerGrid, bool DIV = false>
__global__ void __launch_bounds__(256) growth_uf_kernel(GrowthUF a, Params<double> P,
                                                        GrowthDev g, int LX, int LY, int NP) {
  using T = double;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const int k = n.k;
  const LoopUF<T>& u = a.l;
  auto pl = [&](int i) { return u.D + (size_t)i * NP; };
  const T* phi = u.phi;
  const T* chi = a.chi;
  const T* m = a.m;
  const T* ux = pl(LP_UX);
  const T* uy = pl(LP_UY);
  const T* MU = pl(LP_MU);

  // the phenotype density's increment (go-or-grow.cpp:324-344)
  const T division_m = a.division[k] ? T(1.) : T(0.);
  const T death_m = a.death[k] ? T(1.) : T(0.);
  const T growth_rate = g.alpha * division_m - g.beta * death_m;
  const T chi_eff = DIV ? T(0) : chi[k];
  const T R = DIV ? phi[k] * growth_rate : chi_eff * phi[k] * growth_rate;
  T Dm = T(0);
  if constexpr (!DIV)
    Dm = growth_dm(chi_eff, chi[n.kxp], chi[n.kxm], chi[n.kyp], chi[n.kym], phi[k], phi[n.kxp],
                   phi[n.kxm], phi[n.kyp], phi[n.kym], m[n.kxp], m[n.kxm], m[n.kyp], m[n.kym],
                   ux[n.kxp], ux[n.kxm], uy[n.kyp], uy[n.kym], MU[k], MU[n.kxp], MU[n.kxm],
                   MU[n.kyp], MU[n.kym], R, P, g);

  // Lyotropic::UpdateFieldsAtNode (lyotropic.cpp:249-341)
  CenterVals<T> c;
  c.nn = pl(LP_N)[k];
  c.dn = T(0);
  c.vx = ux[k];
  c.vy = uy[k];
  c.Hxx = pl(LP_HXX)[k];
  c.Hyx = pl(LP_HYX)[k];
  c.dxQxx = pl(LP_DXQXX)[k];
  c.dyQxx = pl(LP_DYQXX)[k];
  c.dxQyx = pl(LP_DXQYX)[k];
  c.dyQyx = pl(LP_DYQYX)[k];
  UFStencil<T, true> s;
  s.mu_c = MU[k];
  s.ux_xp = ux[n.kxp];
  s.ux_xm = ux[n.kxm];
  s.uy_xp = uy[n.kxp];
  s.uy_xm = uy[n.kxm];
  s.uxp_xp = pl(LP_UXP)[n.kxp];
  s.uxp_xm = pl(LP_UXP)[n.kxm];
  s.sxx_xp = pl(LP_SXX)[n.kxp];
  s.sxx_xm = pl(LP_SXX)[n.kxm];
  s.syx_xp = pl(LP_SYX)[n.kxp];
  s.syx_xm = pl(LP_SYX)[n.kxm];
  s.mu_xp = MU[n.kxp];
  s.mu_xm = MU[n.kxm];
  s.ux_y = ynb<T, true>(ux[n.kyp], ux[n.kym]);
  s.uy_y = ynb<T, true>(uy[n.kyp], uy[n.kym]);
  s.uyp_y = ynb<T, true>(pl(LP_UYP)[n.kyp], pl(LP_UYP)[n.kym]);
  s.sxy_y = ynb<T, true>(pl(LP_SXY)[n.kyp], pl(LP_SXY)[n.kym]);
  s.syy_y = ynb<T, true>(pl(LP_SYY)[n.kyp], pl(LP_SYY)[n.kym]);
  s.mu_y = ynb<T, true>(MU[n.kyp], MU[n.kym]);
  T fl[9];
#pragma unroll
  for (int v = 0; v < 9; ++v) fl[v] = u.f[v * NP + k];
  const T Qxx = u.qxx[k], Qyx = u.qyx[k], p = phi[k];
  T Dxx, Dyx, Dp, om[9];
  update_fields<T, true>(c, Qxx, Qyx, p, fl, s, P, Dxx, Dyx, Dp, om);
  // conserve_phi (lyotropic.cpp:297-298; on when neither division nor death
  // can happen, lyotropic-with-division.cpp:31)
  Dp = Dp - (a.conserve ? (*a.countphi - a.totalphi) / T(LX * LY) : T(0));
  T phi_new;
  if constexpr (FIRST) {
    u.qnxx_out[k] = Qxx + T(.5) * Dxx;
    u.qnyx_out[k] = Qyx + T(.5) * Dyx;
    u.phn_out[k] = p + T(.5) * Dp;
    u.qxx_out[k] = Qxx + Dxx;
    u.qyx_out[k] = Qyx + Dyx;
    phi_new = p + Dp;
#pragma unroll
    for (int v = 0; v < 9; ++v) {
      u.fn_out[v * NP + k] = fl[v] + T(.5) * om[v];
      u.f_out[v * NP + k] = fl[v] + om[v];
    }
  } else {
    u.qxx_out[k] = u.qnxx[k] + T(.5) * Dxx;
    u.qyx_out[k] = u.qnyx[k] + T(.5) * Dyx;
    phi_new = u.phn[k] + T(.5) * Dp;
#pragma unroll
    for (int v = 0; v < 9; ++v) u.f_out[v * NP + k] = u.fn[v * NP + k] + T(.5) * om[v];
  }
  // division and death (go-or-grow.cpp:349-350), then the bounded m
  phi_new += R;
  u.phi_out[k] = phi_new;
  if constexpr (DIV) return;
  T mt;
  if constexpr (FIRST) {
    a.mn_out[k] = m[k] + T(.5) * Dm;
    mt = m[k] + Dm;
  } else {
    mt = a.mn[k] + T(.5) * Dm;
  }
  const T upper = phi_new > 0 ? phi_new : T(0.);
  if (mt < 0)
    mt = 0;
  else if (mt > upper)
    mt = upper;
  a.m_out[k] = mt;
  // UpdatePhenotypeQuantities (go-or-grow.cpp:201-209) on the values just stored
  if (a.chi_out) a.chi_out[k] = phi_new > 1e-12 ? mt / phi_new : 0.;
}

// UpdatePhenotypeQuantities (go-or-grow.cpp:201-209): chi = m/phi on the
// domain (0 where phi <= 1e-12).
template <class G = LayerGrid>
__global__ void __launch_bounds__(256) phenotype_kernel(const double* __restrict__ m,
                                                        const double* __restrict__ phi,
                                                        double* __restrict__ chi, int LX, int LY) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const int k = G::at(t, LX, LY).k;
  const double phi_epsilon = 1e-12;
  chi[k] = phi[k] > phi_epsilon ? m[k] / phi[k] : 0.;
}

// ProjectM (go-or-grow.cpp:211-222), after m = phi*chi (:196-197).
template <class G = LayerGrid>
__global__ void __launch_bounds__(256) phenotype_init_kernel(const double* __restrict__ phi,
                                                             const double* __restrict__ chi,
                                                             double* __restrict__ m, int LX,
                                                             int LY) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const int k = G::at(t, LX, LY).k;
  double mk = phi[k] * chi[k];
  const double upper = phi[k] > 0 ? phi[k] : 0.;
  if (mk < 0)
    mk = 0;
  else if (mk > upper)
    mk = upper;
  m[k] = mk;
}

// One smoothing step of the phenotype noise (go-or-grow.cpp:157-162).
template <class G = LayerGrid>
__global__ void __launch_bounds__(256) chi_smooth_kernel(const double* __restrict__ chi,
                                                         double* __restrict__ out, int LX,
                                                         int LY) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const double smooth_rate = 0.2;
  out[n.k] = chi[n.k] +
             smooth_rate * lap5<double, true>(chi[n.k], chi[n.kxp], chi[n.kxm], chi[n.kyp], chi[n.kym]);
}

// RelaxFreeEnergy, first loop (go-or-grow.cpp:53-68): H and MU.
template <class G = LayerGrid>
__global__ void __launch_bounds__(256)
relax_mu_kernel(const double* __restrict__ qxx, const double* __restrict__ qyx,
                const double* __restrict__ phi, Params<double> P, GrowthDev g, int LX, int LY,
                double* __restrict__ hxx, double* __restrict__ hyx, double* __restrict__ mu) {
  using T = double;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const int k = n.k;
  const T Qxx = qxx[k], Qyx = qyx[k], p = phi[k];
  const T q2 = Qxx * Qxx + Qyx * Qyx;
  const T term = g.Snem - q2;
  const T dp_critical = p - g.phi_critical;
  const T mu_compress = dp_critical > 0 ? g.B * dp_critical : T(0.);
  hxx[k] = T(2) * P.CC * term * Qxx +
           P.LL * lap5<T, true>(Qxx, qxx[n.kxp], qxx[n.kxm], qxx[n.kyp], qxx[n.kym]);
  hyx[k] = T(2) * P.CC * term * Qyx +
           P.LL * lap5<T, true>(Qyx, qyx[n.kxp], qyx[n.kxm], qyx[n.kyp], qyx[n.kym]);
  mu[k] = P.AA * p * (T(1) - p) * (T(1) - T(2) * p) + mu_compress -
          P.KK * lap5<T, true>(p, phi[n.kxp], phi[n.kxm], phi[n.kyp], phi[n.kym]);
}

// RelaxFreeEnergy, second loop (go-or-grow.cpp:85-104).
template <class G = LayerGrid>
__global__ void __launch_bounds__(256)
relax_update_kernel(const double* __restrict__ qxx, const double* __restrict__ qyx,
                    const double* __restrict__ phi, const double* __restrict__ hxx,
                    const double* __restrict__ hyx, const double* __restrict__ mu,
                    Params<double> P, double relax_dt, int relax_phi, int relax_Q, int LX, int LY,
                    double* __restrict__ phi_out, double* __restrict__ qnxx,
                    double* __restrict__ qnyx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const LoopNbr n = G::at(t, LX, LY);
  const int k = n.k;
  if (relax_phi)
    phi_out[k] = phi[k] + relax_dt * P.GammaP *
                              lap5<double, true>(mu[k], mu[n.kxp], mu[n.kxm], mu[n.kyp], mu[n.kym]);
  else
    phi_out[k] = phi[k];
  if (relax_Q) {
    qnxx[k] = qxx[k] + relax_dt * P.GammaQ * hxx[k];
    qnyx[k] = qyx[k] + relax_dt * P.GammaQ * hyx[k];
  } else {
    qnxx[k] = qxx[k];
    qnyx[k] = qyx[k];
  }
}

// ResetHydrodynamics (go-or-grow.cpp:121-141): rest equilibrium fe in every
// population buffer and the velocity fields at rest.
template <class G = LayerGrid>
__global__ void __launch_bounds__(256)
reset_hydro_kernel(double* f0, double* f1, double* fn0, double* fn1, double* D, double rho,
                   const double* __restrict__ fe, int LX, int LY, int NP) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= LX * LY) return;
  const int k = G::at(t, LX, LY).k;
  for (int v = 0; v < 9; ++v) {
    const double x = fe[v];
    f0[v * NP + k] = x;
    f1[v * NP + k] = x;
    fn0[v * NP + k] = x;
    fn1[v * NP + k] = x;
  }
  D[(size_t)LP_UX * NP + k] = 0.;
  D[(size_t)LP_UY * NP + k] = 0.;
  D[(size_t)LP_UXP * NP + k] = 0.;
  D[(size_t)LP_UYP * NP + k] = 0.;
  D[(size_t)LP_N * NP + k] = rho;
}

// Marks the disks of squared radius r2 around the centres (x, y pairs): the
// nodes whose periodic distance to a centre is within it
// (lyotropic-with-division.cpp:55-71).  Thread i: centre i / span^2, offset
// i % span^2 within the centre's box of half width R.
__global__ void __launch_bounds__(256) mask_disk_kernel(const int* __restrict__ centres,
                                                        int ncentres, int R, double r2, int LX,
                                                        int LY, unsigned char* __restrict__ mask) {
  const long long span = 2 * R + 1;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ncentres * span * span) return;
  const int c = (int)(i / (span * span));
  const int o = (int)(i % (span * span));
  const int dx = o / (int)span - R, dy = o % (int)span - R;
  const int cx = centres[2 * c], cy = centres[2 * c + 1];
  int x = (cx + dx) % LX, y = (cy + dy) % LY;
  if (x < 0) x += LX;
  if (y < 0) y += LY;
  // wrap(diff(a, b), L) = min(|a-b| mod L, L - |a-b| mod L)
  const unsigned adx = (unsigned)(x > cx ? x - cx : cx - x) % (unsigned)LX;
  const unsigned ady = (unsigned)(y > cy ? y - cy : cy - y) % (unsigned)LY;
  const unsigned wx = adx < LX - adx ? adx : LX - adx;
  const unsigned wy = ady < LY - ady ? ady : LY - ady;
  if ((double)wy * (double)wy + (double)wx * (double)wx <= r2)
    mask[LayerGrid::at(y + LY * x, LX, LY).k] = 1;
}

}  // namespace lyo
