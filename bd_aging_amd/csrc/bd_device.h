// bd_device.h -- device-side numerics of the Brownian-dynamics engine
// (gfx950).  Every function here is evaluated with IEEE-754 double
// operations only (+ - * / sqrt fma floor rint and integer ops) and the
// translation unit is compiled with -ffp-contract=off, so each result is a
// fixed function of its inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bd_abi.h"

namespace bd {

// 2^(1/6), the WCA cutoff factor (CUTOFF_FAC, encabezados.h:26, in FP64).
constexpr double kWcaCut = 0x1.1f59ac3c7d6c0p+0;

// rannew64 constants (rannew64.cu:52-53,58,73).
constexpr int kRingLen = 250;
constexpr int kLagGap = 147;  // np - nq (mod 250): lags 250 and 103
constexpr uint64_t kLcgA = 2862933555777941757ULL;
constexpr uint64_t kLcgC = 3037000493ULL;
constexpr uint64_t kSeedA = 6364136223846793005ULL;
constexpr uint64_t kSeedC = 1442695040888963407ULL;

// Lane seed: splitmix64 finaliser (Steele, Lea & Flood 2014) of
// seed + (lane+1)*golden, folded to a non-negative idum.
__host__ __device__ inline uint64_t lane_seed(uint64_t seed, uint64_t lane) {
  uint64_t z = seed + (lane + 1ULL) * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z = z ^ (z >> 31);
  return z >> 1;
}

// ln(x) for a positive normal x: x = 2^e m with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s), s = (m-1)/(m+1), series through s^21.
__device__ inline double log_pos(double x) {
  uint64_t b = (uint64_t)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m = __longlong_as_double((long long)b);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e = e + 1;
  }
  double f = m - 1.0;
  double s = f / (m + 1.0);
  double z = s * s;
  double p = 0x1.8618618618618p-5;
  p = fma(p, z, 0x1.af286bca1af28p-5);
  p = fma(p, z, 0x1.e1e1e1e1e1e1ep-5);
  p = fma(p, z, 0x1.1111111111111p-4);
  p = fma(p, z, 0x1.3b13b13b13b14p-4);
  p = fma(p, z, 0x1.745d1745d1746p-4);
  p = fma(p, z, 0x1.c71c71c71c71cp-4);
  p = fma(p, z, 0x1.2492492492492p-3);
  p = fma(p, z, 0x1.999999999999ap-3);
  p = fma(p, z, 0x1.5555555555555p-2);
  double t = (s * z) * p;
  double lm = 2.0 * (s + t);
  double de = (double)e;
  return de * 0x1.62e42fefa3800p-1 + (de * 0x1.ef35793c76730p-45 + lm);
}

// sin and cos of 2*pi*u for u in [0, 1]: quadrant from rint(4u) (exact
// reduction), Taylor series on [-pi/4, pi/4].
__device__ inline void sincos_turn(double u, double& sn, double& cs) {
  double t = 4.0 * u;
  double k = rint(t);
  double f = t - k;
  double x = fma(f, 0x1.921fb54442d18p+0, f * 0x1.1a62633145c07p-54);
  double z = x * x;
  double ps = 0x1.952c77030ad4ap-49;
  ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
  ps = fma(ps, z, 0x1.6124613a86d09p-33);
  ps = fma(ps, z, -0x1.ae64567f544e4p-26);
  ps = fma(ps, z, 0x1.71de3a556c734p-19);
  ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
  ps = fma(ps, z, 0x1.1111111111111p-7);
  ps = fma(ps, z, -0x1.5555555555555p-3);
  double sx = fma(x * z, ps, x);
  double pc = -0x1.6827863b97d97p-53;
  pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);
  pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
  pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
  pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
  pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
  pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
  pc = fma(pc, z, 0x1.5555555555555p-5);
  double cx = fma(z * z, pc, fma(-0.5, z, 1.0));
  int q = ((int)k) & 3;
  if (q == 0) {
    sn = sx;
    cs = cx;
  } else if (q == 1) {
    sn = cx;
    cs = -sx;
  } else if (q == 2) {
    sn = -sx;
    cs = -cx;
  } else {
    sn = -cx;
    cs = sx;
  }
}

// Three standard normals from four rannew64 words (Box-Muller on the 53
// high bits; the fourth normal is discarded).
__device__ inline void gauss3(uint64_t d0, uint64_t d1, uint64_t d2, uint64_t d3, double& x,
                              double& y, double& z) {
  const double two_m53 = 0x1.0p-53;
  double u1 = (double)((d0 >> 11) + 1ULL) * two_m53;
  double u2 = (double)(d1 >> 11) * two_m53;
  double u3 = (double)((d2 >> 11) + 1ULL) * two_m53;
  double u4 = (double)(d3 >> 11) * two_m53;
  double r1 = sqrt(-2.0 * log_pos(u1));
  double r2 = sqrt(-2.0 * log_pos(u3));
  double s, c;
  sincos_turn(u2, s, c);
  x = r1 * c;
  y = r1 * s;
  sincos_turn(u4, s, c);
  z = r2 * c;
}

// Pair potentials (potentials.cu, FP64): e is the pair's energy term, added
// to the running potential as pot = pot + e; nf is dU/dr.  q holds the
// per-system constants: kappa, amp_t = pot_amp * pot_inv_t, inv_t, sigma_pol,
// phi_pol.
//   WCA          potentials.cu:54-64
//   WCA_MODIFIED potentials.cu:66-76
//   WCA49_50     potentials.cu:9-23 (s^6 as (s^3)^2, as the host form does)
//   WCA49_50_AO  potentials.cu:27-52 (s^6 as six factors; the depletion
//                branch beyond r = sigma)
// EPS = 1 and 1 * a == a, so the reference's EPS*AA*invT is amp_t and
// -EPS*AA*invT is -amp_t bit for bit.
__device__ inline void repulsion_49_50(double amp_t, double inv_t, double sor, double sor6,
                                       double dist_inv, double& e, double& nf) {
  double sor12 = sor6 * sor6;
  double sor24 = sor12 * sor12;
  double sor48 = sor24 * sor24;
  double sor49 = sor48 * sor;
  double sor50 = sor49 * sor;
  e = amp_t * (sor50 - sor49) + inv_t;
  nf = -amp_t * (50.0 * sor50 - 49.0 * sor48 * sor) * dist_inv;
}

template <int KIND, class P>
__device__ inline void pair_energy(const P& q, double sigma, double dist_inv, double& e,
                                   double& nf) {
  if (KIND == BD_POT_HARMONIC) {
    double r = 1.0 / dist_inv;
    double delta = sigma - r;
    e = (0.5 * q.kappa) * delta * delta;
    nf = -q.kappa * delta;
  } else if (KIND == BD_POT_WCA49_50) {
    double sor = dist_inv * sigma;
    double sor6 = sor * sor * sor;
    sor6 = sor6 * sor6;
    repulsion_49_50(q.amp_t, q.inv_t, sor, sor6, dist_inv, e, nf);
  } else if (KIND == BD_POT_WCA49_50_AO) {
    double size_ratio = sigma / q.sigma_pol;
    double sor = dist_inv * sigma;
    double ros = 1.0 / sor;
    if (ros <= 1.0) {
      double sor6 = sor * sor * sor * sor * sor * sor;
      repulsion_49_50(q.amp_t, q.inv_t, sor, sor6, dist_inv, e, nf);
    } else {
      double ros3 = ros * ros * ros;
      double eta_1 = size_ratio + 1.0;
      double eta2 = eta_1 * eta_1;
      e = -q.phi_pol * (eta2 * eta_1 - 1.5 * eta2 * ros + 0.5 * ros3);
      nf = q.phi_pol * (1.5 * eta2 * ros - 1.5 * ros3) * dist_inv;
    }
  } else {
    double sor = dist_inv * sigma;
    double sor6 = sor * sor * sor;
    sor6 = sor6 * sor6;
    double sor12 = sor6 * sor6;
    if (KIND == BD_POT_WCA) {
      e = 4.0 * (sor12 - sor6) + 1.0;
      nf = -4.0 * (12.0 * sor12 - 6.0 * sor6) * dist_inv;
    } else {
      e = (sor12 - 2.0 * sor6) + 1.0;
      nf = -12.0 * (sor12 - sor6) * dist_inv;
    }
  }
}

template <int KIND, class P>
__device__ inline void potential(const P& q, double sigma, double dist_inv, double& pot,
                                 double& nf) {
  double e;
  pair_energy<KIND>(q, sigma, dist_inv, e, nf);
  pot = pot + e;
}

}  // namespace bd
