/*
 * bd_oracle.c -- FP64 single-threaded CPU restatement of the reference
 * Brownian-dynamics path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is
 * compared against and the CPU baseline timed beside it; the product
 * (libbd_hip.so, bd_aging_self) never links or calls it.
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math so that every a*b+c is two
 * roundings, as in the engine.  Paths in citations are relative to the
 * reference tree.
 *
 * Pinning (what checks each part, besides the GPU parity tests):
 *   rannew64            golden vectors of the compiled reference rannew64.cu
 *   orc_log / sincos    libm to a few ulp; Box-Muller moments and KS test
 *   geometry            reference cell counts for the reference parameters
 *   cell locate         brute-force cell index + ascending order invariant
 *   forces              all-pairs brute force, Newton 3 (sum F = 0), the
 *                       identity wca_modified(s*2^(1/6)) == wca(s), closed form
 *   update              closed form with zero noise; free-diffusion MSD = 6Dt
 *   temperature         closed forms at the protocol break points
 *   sampling indexes    float32 emulation fixture of main.cu:14-53
 *   g(r)                brute-force counts; total = ordered pairs in range
 *   observables         free diffusion F_s = exp(-k^2 D t); closed forms
 *   cv3d                exact recovery of a bin-wise constant radial force;
 *                       pseudo-inverse against numpy.linalg.pinv
 */
#include "bd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_NP 250 /* rannew64.cu:52 */
#define ORC_NQ 103 /* rannew64.cu:53 */

/* 2^(1/6): WCA cutoff factor (CUTOFF_FAC, encabezados.h:26, in FP64) */
static const double WCA_CUT = 0x1.1f59ac3c7d6c0p+0;
/* PI as the reference defines it: (double)3.14159266f (encabezados.h:25) */
static const double PI_REF = 3.1415927410125732;

/* ======================================================================
 * rannew64 (rannew64.cu:55-126), one stream per orc_rn64.
 * The reference keeps np/nq in statics that a re-seed does not reset
 * (rannew64.cu:63-64); here every seed starts a fresh stream, which is what
 * the reference does on its first call in a fresh process.
 * ==================================================================== */
void orc_rn64_seed(orc_rn64* s, int64_t idum) {
  /* rannew64.cu:77-81: a negative idum is negated */
  if (idum < 0) idum = -idum;
  uint64_t udum = (uint64_t)idum;
  const uint64_t uaa = 6364136223846793005ULL, ucc = 1442695040888963407ULL;
  /* rannew64.cu:86-91: 1000 warm-up LCG draws, then fill the ring */
  for (int i = 0; i < 1000; i++) udum = udum * uaa + ucc;
  for (int i = 0; i < ORC_NP; i++) {
    udum = udum * uaa + ucc;
    s->ring[i] = udum;
  }
  s->lcg = udum; /* rannew64.cu:95 */
  s->np = ORC_NP - 1;
  s->nq = ORC_NQ - 1;
}

uint64_t orc_rn64_u64(orc_rn64* s) {
  /* rannew64.cu:101-106: lagged Fibonacci x(n) = x(n-250) - x(n-103) */
  uint64_t uux = s->ring[s->np] - s->ring[s->nq];
  s->ring[s->np] = uux;
  s->np--;
  s->nq--;
  if (s->np < 0) s->np = ORC_NP - 1;
  if (s->nq < 0) s->nq = ORC_NP - 1;
  /* rannew64.cu:110: y(n) = a y(n-1) + c */
  s->lcg = 2862933555777941757ULL * s->lcg + 3037000493ULL;
  /* rannew64.cu:114 */
  return uux - s->lcg;
}

double orc_rn64_double(orc_rn64* s) {
  /* rannew64.cu:59,94,125: omm = 1/18446744073709551615. == 2^-64 */
  const double omm = 1.0 / 18446744073709551615.0;
  return omm * (double)orc_rn64_u64(s);
}

/* Lane seed: splitmix64 finaliser (Steele, Lea & Flood 2014) of
 * seed + (lane+1)*golden, folded to a non-negative idum. */
uint64_t orc_lane_seed(uint64_t seed, uint64_t lane) {
  uint64_t z = seed + (lane + 1ULL) * 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z = z ^ (z >> 31);
  return z >> 1;
}

/* ======================================================================
 * Gaussian transform.  Box & Muller (1958) on the 53 high bits of two
 * rannew64 words.  log and sincos are written here from their series so
 * that the CPU and the GPU evaluate the same IEEE operations:
 *   log(m) = 2 atanh((m-1)/(m+1)), m in [sqrt(1/2), sqrt(2)), ten terms;
 *   sin/cos(2 pi u) by quadrant reduction of 4u (exact) and Taylor series
 *   on [-pi/4, pi/4] through x^17 / x^18.
 * ==================================================================== */
static const double LN2_HI = 0x1.62e42fefa3800p-1;
static const double LN2_LO = 0x1.ef35793c76730p-45;
static const double SQRT2 = 0x1.6a09e667f3bcdp+0;
static const double PIO2_HI = 0x1.921fb54442d18p+0;
static const double PIO2_LO = 0x1.1a62633145c07p-54;

double orc_log(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  memcpy(&m, &b, 8);
  if (m > SQRT2) {
    m = m * 0.5;
    e = e + 1;
  }
  double f = m - 1.0;
  double s = f / (m + 1.0);
  double z = s * s;
  double p = 0x1.8618618618618p-5;    /* 1/21 */
  p = fma(p, z, 0x1.af286bca1af28p-5); /* 1/19 */
  p = fma(p, z, 0x1.e1e1e1e1e1e1ep-5); /* 1/17 */
  p = fma(p, z, 0x1.1111111111111p-4); /* 1/15 */
  p = fma(p, z, 0x1.3b13b13b13b14p-4); /* 1/13 */
  p = fma(p, z, 0x1.745d1745d1746p-4); /* 1/11 */
  p = fma(p, z, 0x1.c71c71c71c71cp-4); /* 1/9 */
  p = fma(p, z, 0x1.2492492492492p-3); /* 1/7 */
  p = fma(p, z, 0x1.999999999999ap-3); /* 1/5 */
  p = fma(p, z, 0x1.5555555555555p-2); /* 1/3 */
  double t = (s * z) * p;
  double lm = 2.0 * (s + t);
  double de = (double)e;
  return de * LN2_HI + (de * LN2_LO + lm);
}

void orc_sincos_turn(double u, double* sn, double* cs) {
  double t = 4.0 * u;
  double k = rint(t);
  double f = t - k;
  double x = fma(f, PIO2_HI, f * PIO2_LO);
  double z = x * x;
  double ps = 0x1.952c77030ad4ap-49;     /* 1/17! */
  ps = fma(ps, z, -0x1.ae7f3e733b81fp-41); /* -1/15! */
  ps = fma(ps, z, 0x1.6124613a86d09p-33);  /* 1/13! */
  ps = fma(ps, z, -0x1.ae64567f544e4p-26); /* -1/11! */
  ps = fma(ps, z, 0x1.71de3a556c734p-19);  /* 1/9! */
  ps = fma(ps, z, -0x1.a01a01a01a01ap-13); /* -1/7! */
  ps = fma(ps, z, 0x1.1111111111111p-7);   /* 1/5! */
  ps = fma(ps, z, -0x1.5555555555555p-3);  /* -1/3! */
  double sx = fma(x * z, ps, x);
  double pc = -0x1.6827863b97d97p-53;    /* -1/18! */
  pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);  /* 1/16! */
  pc = fma(pc, z, -0x1.93974a8c07c9dp-37); /* -1/14! */
  pc = fma(pc, z, 0x1.1eed8eff8d898p-29);  /* 1/12! */
  pc = fma(pc, z, -0x1.27e4fb7789f5cp-22); /* -1/10! */
  pc = fma(pc, z, 0x1.a01a01a01a01ap-16);  /* 1/8! */
  pc = fma(pc, z, -0x1.6c16c16c16c17p-10); /* -1/6! */
  pc = fma(pc, z, 0x1.5555555555555p-5);   /* 1/4! */
  double cx = fma(z * z, pc, fma(-0.5, z, 1.0));
  int q = ((int)k) & 3;
  if (q == 0) {
    *sn = sx;
    *cs = cx;
  } else if (q == 1) {
    *sn = cx;
    *cs = -sx;
  } else if (q == 2) {
    *sn = -sx;
    *cs = -cx;
  } else {
    *sn = -cx;
    *cs = sx;
  }
}

void orc_gauss3(const uint64_t d[4], double out[3]) {
  const double two_m53 = 0x1.0p-53;
  double u1 = (double)((d[0] >> 11) + 1ULL) * two_m53; /* (0, 1] */
  double u2 = (double)(d[1] >> 11) * two_m53;          /* [0, 1) */
  double u3 = (double)((d[2] >> 11) + 1ULL) * two_m53;
  double u4 = (double)(d[3] >> 11) * two_m53;
  double r1 = sqrt(-2.0 * orc_log(u1));
  double r2 = sqrt(-2.0 * orc_log(u3));
  double s, c;
  orc_sincos_turn(u2, &s, &c);
  out[0] = r1 * c;
  out[1] = r1 * s;
  orc_sincos_turn(u4, &s, &c);
  out[2] = r2 * c;
}

/* ======================================================================
 * Geometry (main.cu:259-298).  The reference sizes cells at 1.05 sigma for
 * its active cutoff sigma (wca_modified); here the cutoff rc of the chosen
 * potential replaces sigma, and nrange_sb uses the same formula as nrange_bs
 * (the reference hard-codes 1, which the formula gives when sigma_big >=
 * sigma_sml).  A grid narrower than 2*nrange+1 cells would visit a cell
 * twice, so it is rejected.
 * ==================================================================== */
/* Pair cutoff: 2^(1/6) sigma (WCA); (50/49) sigma (49-50, the commented
 * cutoff=BB*sigma of get_forces.cu:47-49); sigma (1 + sigma/sigma_pol) for
 * 49-50 + AO, the distance at which the depletion term of
 * potentials.cu:43-50 and its force reach zero; sigma otherwise. */
static double pair_cut(const bd_params* p, double sigma) {
  if (p->potential == BD_POT_WCA) return WCA_CUT * sigma;
  if (p->potential == BD_POT_WCA49_50) return (50.0 / 49.0) * sigma;
  if (p->potential == BD_POT_WCA49_50_AO) return sigma * (1.0 + sigma / p->sigma_pol);
  return 1.0 * sigma;
}

int orc_geometry(const bd_params* p, bd_geometry* g) {
  if (!p || !g) return BD_E_ARG;
  if (!(p->side > 0.0) || !(p->sigma_big > 0.0) || !(p->sigma_sml > 0.0)) return BD_E_ARG;
  if (p->n_big < 1 || p->n_sml < 1) return BD_E_ARG;
  if (p->ntags_big < 1 || p->ntags_big > 32 || p->ntags_sml < 1 || p->ntags_sml > 32)
    return BD_E_ARG;
  if (p->potential < BD_POT_WCA || p->potential > BD_POT_WCA49_50_AO) return BD_E_ARG;
  if ((p->potential == BD_POT_WCA49_50 || p->potential == BD_POT_WCA49_50_AO) &&
      !(p->pot_amp > 0.0 && p->pot_inv_t > 0.0 && isfinite(p->pot_amp * p->pot_inv_t)))
    return BD_E_ARG;
  if (p->potential == BD_POT_WCA49_50_AO &&
      !(p->sigma_pol > 0.0 && p->phi_pol >= 0.0 && isfinite(p->phi_pol)))
    return BD_E_ARG;
  double sbs = 0.5 * (p->sigma_big + p->sigma_sml);
  g->rc_bb = pair_cut(p, p->sigma_big);
  g->rc_ss = pair_cut(p, p->sigma_sml);
  g->rc_bs = pair_cut(p, sbs);
  if (!isfinite(g->rc_bb) || !isfinite(g->rc_ss) || !isfinite(g->rc_bs)) return BD_E_ARG;
  g->ncell_big = (int32_t)(p->side / (1.05 * g->rc_bb));
  g->ncell_sml = (int32_t)(p->side / (1.05 * g->rc_ss));
  if (g->ncell_big < 1 || g->ncell_sml < 1) return BD_E_ARG;
  g->cell_side_big = p->side / (double)g->ncell_big;
  g->cell_side_sml = p->side / (double)g->ncell_sml;
  g->nrange_bb = 1 + (int32_t)(g->rc_bb / g->cell_side_big);
  g->nrange_ss = 1 + (int32_t)(g->rc_ss / g->cell_side_sml);
  g->nrange_bs = 1 + (int32_t)(g->rc_bs / g->cell_side_sml); /* main.cu:281-283 */
  g->nrange_sb = 1 + (int32_t)(g->rc_bs / g->cell_side_big);
  int rb = g->nrange_bb > g->nrange_sb ? g->nrange_bb : g->nrange_sb;
  int rs = g->nrange_ss > g->nrange_bs ? g->nrange_ss : g->nrange_bs;
  if (g->ncell_big < 2 * rb + 1 || g->ncell_sml < 2 * rs + 1) return BD_E_ARG;
  return BD_OK;
}

/* ======================================================================
 * Temperature protocols (set_temp.cu:7-23).
 * Deviation: with tf <= t0 (instant quench, run-fself.sh:30-31) the
 * reference divides 0 by 0 at t == t0 and returns NaN; here t >= t0 gives
 * Tf.  The sine uses an exact 2*pi turn (sin(2 pi u) of the Gaussian
 * transform) instead of 2*3.14159266f.
 * ==================================================================== */
double orc_temp_linear(double T0, double Tf, double t0, double tf, double t) {
  if (t < t0) return T0;
  if (t > tf || !(tf > t0)) return Tf;
  double m = (Tf - T0) / (tf - t0);
  return T0 + m * (t - t0);
}

double orc_temp_sine(double T0, double Tf, double t0, double tf, double period, double t) {
  if (t < t0) return T0;
  if (t > tf) return Tf;
  double a = 0.5 * (T0 + Tf);
  double b = 0.5 * (Tf - T0);
  double u = t / period - 0.25;
  u = u - floor(u);
  double s, c;
  orc_sincos_turn(u, &s, &c);
  return a + b * s;
}

/* Temperature used by the update at step ni: pars.temp_set starts at T0
 * (main.cu:155) and after the update of step ni-1 becomes
 * T(dt*(1+(ni-1))) (main.cu:520-524). */
double orc_temperature(const bd_params* p, int64_t ni) {
  if (ni == p->first_step) return p->T0;
  double t = p->dt * (1.0 + (double)(ni - 1));
  if (p->temp_protocol == BD_TEMP_SINE)
    return orc_temp_sine(p->T0, p->Tf, p->t0, p->tf, p->period, t);
  return orc_temp_linear(p->T0, p->Tf, p->t0, p->tf, t);
}

/* main.cu:14-53 in the reference's float arithmetic. */
int32_t orc_sampling_indexes(float minimal_decade, float run_time, float dt,
                             int32_t per_decade, int32_t* out, int32_t maxn) {
  int last_power = (int)roundf(log10f(run_time));
  int initial_power = (int)roundf(log10f(minimal_decade));
  int decades = last_power - initial_power;
  if (per_decade < 1 || decades < 0) return -1;
  int total = decades * per_decade;
  if (total + 1 > maxn) return -1;
  float* values = (float*)malloc(sizeof(float) * (size_t)(per_decade + 1));
  float r = powf(10.0f, 1.0f / (float)per_decade);
  float x = 1.0f;
  for (int i = 0; i <= per_decade; ++i) {
    values[i] = x;
    x *= r;
  }
  int cnt = 0;
  out[0] = 0;
  for (int ipow = initial_power; ipow < last_power; ++ipow) {
    float power = (float)pow(10.0, (double)ipow);
    for (int j = 0; j < per_decade; ++j) {
      float scaled_time = values[j] * power;
      int index = (int)roundf(scaled_time / dt);
      cnt++;
      out[cnt] = index;
    }
  }
  free(values);
  return total + 1;
}

/* ======================================================================
 * Pair potentials (potentials.cu) in FP64.  normal_force is dU/dr; the
 * energy term is added as potential = potential + term.
 * HARMONIC (overlap removal, no reference counterpart):
 *   U = kappa/2 (sigma - r)^2 for r < sigma.
 * WCA49_50 follows potential_wca49_50_hst (potentials.cu:9-23) and
 * WCA49_50_AO potential_wca49_50_AO_hst (potentials.cu:27-52) with
 * AA*invT = pot_amp*pot_inv_t; EPS = 1, so EPS*AA*invT is that product and
 * -EPS*AA*invT its negation, bit for bit.  The two host forms build s^6
 * differently ((s^3)^2 and six factors); each is kept.
 * ==================================================================== */
static void repulsion_49_50(double ai, double inv_t, double sor, double sor6, double dist_inv,
                            double* potential, double* normal_force) {
  double sor12 = sor6 * sor6;
  double sor24 = sor12 * sor12;
  double sor48 = sor24 * sor24;
  double sor50 = sor48 * sor * sor;
  *potential = *potential + (ai * (sor50 - sor48 * sor) + inv_t);
  *normal_force = -ai * (50.0 * sor50 - 49.0 * sor48 * sor) * dist_inv;
}

void orc_potential(const bd_params* p, double sigma, double dist_inv, double* potential,
                   double* normal_force) {
  int32_t kind = p->potential;
  if (kind == BD_POT_HARMONIC) {
    double r = 1.0 / dist_inv;
    double delta = sigma - r;
    *potential = *potential + (0.5 * p->kappa) * delta * delta;
    *normal_force = -p->kappa * delta;
    return;
  }
  double sor = dist_inv * sigma;
  if (kind == BD_POT_WCA49_50) {
    double sor6 = sor * sor * sor;
    sor6 = sor6 * sor6;
    repulsion_49_50(p->pot_amp * p->pot_inv_t, p->pot_inv_t, sor, sor6, dist_inv, potential,
                    normal_force);
    return;
  }
  if (kind == BD_POT_WCA49_50_AO) {
    double size_ratio = sigma / p->sigma_pol;
    double ros = 1.0 / sor;
    if (ros <= 1.0) {
      double sor6 = sor * sor * sor * sor * sor * sor;
      repulsion_49_50(p->pot_amp * p->pot_inv_t, p->pot_inv_t, sor, sor6, dist_inv, potential,
                      normal_force);
    } else {
      double ros3 = ros * ros * ros;
      double eta_1 = size_ratio + 1.0;
      double eta2 = eta_1 * eta_1;
      *potential = *potential + (-p->phi_pol * (eta2 * eta_1 - 1.5 * eta2 * ros + 0.5 * ros3));
      *normal_force = p->phi_pol * (1.5 * eta2 * ros - 1.5 * ros3) * dist_inv;
    }
    return;
  }
  double sor6 = sor * sor * sor;
  sor6 = sor6 * sor6;
  double sor12 = sor6 * sor6;
  if (kind == BD_POT_WCA) {
    /* potentials.cu:54-64 */
    *potential = *potential + (4.0 * (sor12 - sor6) + 1.0);
    *normal_force = -4.0 * (12.0 * sor12 - 6.0 * sor6) * dist_inv;
  } else {
    /* potentials.cu:66-76 */
    *potential = *potential + ((sor12 - 2.0 * sor6) + 1.0);
    *normal_force = -12.0 * (sor12 - sor6) * dist_inv;
  }
}

/* ======================================================================
 * System state.
 * ==================================================================== */
struct orc_sys {
  bd_params p;
  bd_geometry g;
  int32_t n, nb, ns;
  double *pos, *raw, *ini, *f, *vir, *pot; /* N x 3 / N, big first */
  int32_t* nocup[2];
  int32_t* cell[2];
  int32_t ncell3[2];
  int32_t ntags[2];
  int32_t lanes;
  int32_t per_lane; /* ceil(N / lanes) particles served by a lane per step */
  orc_rn64* rng;
  int64_t ni;
  double temp;
};

orc_sys* orc_create(const bd_params* p) {
  bd_geometry g;
  if (orc_geometry(p, &g) != BD_OK) return NULL;
  orc_sys* s = (orc_sys*)calloc(1, sizeof(orc_sys));
  s->p = *p;
  s->g = g;
  s->nb = p->n_big;
  s->ns = p->n_sml;
  s->n = p->n_big + p->n_sml;
  size_t n3 = (size_t)s->n * 3;
  s->pos = (double*)calloc(n3, sizeof(double));
  s->raw = (double*)calloc(n3, sizeof(double));
  s->ini = (double*)calloc(n3, sizeof(double));
  s->f = (double*)calloc(n3, sizeof(double));
  s->vir = (double*)calloc((size_t)s->n, sizeof(double));
  s->pot = (double*)calloc((size_t)s->n, sizeof(double));
  int32_t nc[2] = {g.ncell_big, g.ncell_sml};
  s->ntags[0] = p->ntags_big;
  s->ntags[1] = p->ntags_sml;
  for (int k = 0; k < 2; k++) {
    s->ncell3[k] = nc[k] * nc[k] * nc[k];
    s->nocup[k] = (int32_t*)calloc((size_t)s->ncell3[k], sizeof(int32_t));
    s->cell[k] = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->ncell3[k] * (size_t)s->ntags[k]);
    for (size_t i = 0; i < (size_t)s->ncell3[k] * (size_t)s->ntags[k]; i++) s->cell[k][i] = -1;
  }
  s->lanes = p->rng_lanes > 0 ? p->rng_lanes : s->n;
  if (s->lanes > s->n) s->lanes = s->n;
  s->per_lane = (s->n + s->lanes - 1) / s->lanes;
  s->rng = (orc_rn64*)malloc(sizeof(orc_rn64) * (size_t)s->lanes);
  for (int32_t l = 0; l < s->lanes; l++)
    orc_rn64_seed(&s->rng[l], (int64_t)orc_lane_seed(p->seed, (uint64_t)l));
  s->ni = p->first_step;
  s->temp = p->T0;
  return s;
}

void orc_destroy(orc_sys* s) {
  if (!s) return;
  free(s->pos);
  free(s->raw);
  free(s->ini);
  free(s->f);
  free(s->vir);
  free(s->pot);
  for (int k = 0; k < 2; k++) {
    free(s->nocup[k]);
    free(s->cell[k]);
  }
  free(s->rng);
  free(s);
}

int orc_get_geometry(const orc_sys* s, bd_geometry* g) {
  *g = s->g;
  return BD_OK;
}

int orc_set_positions(orc_sys* s, const double* xyz) {
  size_t n3 = (size_t)s->n * 3;
  memcpy(s->pos, xyz, n3 * sizeof(double));
  memcpy(s->raw, xyz, n3 * sizeof(double));
  memcpy(s->ini, xyz, n3 * sizeof(double));
  return BD_OK;
}

int orc_get_positions(const orc_sys* s, double* xyz) {
  memcpy(xyz, s->pos, (size_t)s->n * 3 * sizeof(double));
  return BD_OK;
}

int orc_get_raw_positions(const orc_sys* s, double* xyz) {
  memcpy(xyz, s->raw, (size_t)s->n * 3 * sizeof(double));
  return BD_OK;
}

int orc_set_forces(orc_sys* s, const double* f) {
  memcpy(s->f, f, (size_t)s->n * 3 * sizeof(double));
  return BD_OK;
}

int orc_get_forces(const orc_sys* s, double* f, double* vir, double* pot) {
  if (f) memcpy(f, s->f, (size_t)s->n * 3 * sizeof(double));
  if (vir) memcpy(vir, s->vir, (size_t)s->n * sizeof(double));
  if (pot) memcpy(pot, s->pot, (size_t)s->n * sizeof(double));
  return BD_OK;
}

/* main.cu:497-506 */
int orc_mark_origin(orc_sys* s) {
  size_t n3 = (size_t)s->n * 3;
  memcpy(s->raw, s->pos, n3 * sizeof(double));
  memcpy(s->ini, s->pos, n3 * sizeof(double));
  return BD_OK;
}

int64_t orc_get_step(const orc_sys* s) { return s->ni; }
double orc_get_temperature(const orc_sys* s) { return s->temp; }

/* ----------------------------------------------------------------------
 * cell_locate_hst (cell_locate.cu:46-104) after set_vec_int_hst
 * (cell_locate.cu:7-16, main.cu:536-541): ascending particle order within a
 * cell.  A cell with >= ntags particles is an overflow (main.cu:655-659);
 * slots beyond ntags are not written.
 * -------------------------------------------------------------------- */
static void cell_of(double inv, int32_t ncell, const double* r, int32_t* ii, int32_t* jj,
                    int32_t* kk) {
  *ii = (int32_t)(inv * r[0]);
  *jj = (int32_t)(inv * r[1]);
  *kk = (int32_t)(inv * r[2]);
  if (*ii == ncell) (*ii)--;
  if (*jj == ncell) (*jj)--;
  if (*kk == ncell) (*kk)--;
}

int orc_cell_locate(orc_sys* s) {
  int rc = BD_OK;
  for (int k = 0; k < 2; k++) {
    int32_t ncell = k == 0 ? s->g.ncell_big : s->g.ncell_sml;
    double inv = 1.0 / (k == 0 ? s->g.cell_side_big : s->g.cell_side_sml);
    int32_t ntags = s->ntags[k];
    int32_t first = k == 0 ? 0 : s->nb;
    int32_t cnt = k == 0 ? s->nb : s->ns;
    memset(s->nocup[k], 0, sizeof(int32_t) * (size_t)s->ncell3[k]);
    for (size_t i = 0; i < (size_t)s->ncell3[k] * (size_t)ntags; i++) s->cell[k][i] = -1;
    for (int32_t mm = 0; mm < cnt; mm++) {
      int32_t ii, jj, kk;
      cell_of(inv, ncell, &s->pos[3 * (size_t)(first + mm)], &ii, &jj, &kk);
      if (ii < 0 || jj < 0 || kk < 0 || ii >= ncell || jj >= ncell || kk >= ncell)
        return BD_E_NAN;
      int32_t ci = ii + ncell * (jj + ncell * kk);
      int32_t shift = s->nocup[k][ci];
      if (shift < ntags) s->cell[k][(size_t)ci * ntags + shift] = mm;
      s->nocup[k][ci]++;
      if (s->nocup[k][ci] >= ntags) rc = BD_E_CELL_OVERFLOW;
    }
  }
  return rc;
}

int orc_get_cells(const orc_sys* s, int32_t species, int32_t* nocup, int32_t* cell_vec) {
  if (species < 0 || species > 1) return BD_E_ARG;
  memcpy(nocup, s->nocup[species], sizeof(int32_t) * (size_t)s->ncell3[species]);
  memcpy(cell_vec, s->cell[species],
         sizeof(int32_t) * (size_t)s->ncell3[species] * (size_t)s->ntags[species]);
  return BD_OK;
}

/* ----------------------------------------------------------------------
 * get_forces_same_hst (get_forces.cu:21-145) followed by
 * get_forces_diff_hst (get_forces.cu:149-277) for one particle.  The
 * accumulators start at zero (set_vec_float*_hst, main.cu:556-561).  When
 * nbr is not NULL the interacting neighbours are recorded in traversal order.
 * -------------------------------------------------------------------- */
static void scan_grid(const orc_sys* s, int32_t m, int32_t grid, int32_t nrange, double sigma,
                      double rc2, int skip_self, double* fx, double* fy, double* fz,
                      double* vir, double* pot, int32_t* nbr, int32_t* nn_count,
                      int32_t maxn) {
  int32_t ncell = grid == 0 ? s->g.ncell_big : s->g.ncell_sml;
  double inv = 1.0 / (grid == 0 ? s->g.cell_side_big : s->g.cell_side_sml);
  int32_t ntags = s->ntags[grid];
  int32_t first = grid == 0 ? 0 : s->nb;
  double side = s->p.side;
  double side_inv = 1.0 / side;
  const double* rm = &s->pos[3 * (size_t)m];
  int32_t ii, jj, kk;
  cell_of(inv, ncell, rm, &ii, &jj, &kk);
  int32_t mloc = m - (m < s->nb ? 0 : s->nb);
  for (int32_t i_del = -nrange; i_del <= nrange; i_del++)
    for (int32_t j_del = -nrange; j_del <= nrange; j_del++)
      for (int32_t k_del = -nrange; k_del <= nrange; k_del++) {
        int32_t iip = (ii + i_del + ncell) % ncell;
        int32_t jjp = (jj + j_del + ncell) % ncell;
        int32_t kkp = (kk + k_del + ncell) % ncell;
        int32_t ci = iip + ncell * (jjp + ncell * kkp);
        int32_t nocup = s->nocup[grid][ci];
        if (nocup > ntags) nocup = ntags;
        for (int32_t t = 0; t < nocup; t++) {
          int32_t nn = s->cell[grid][(size_t)ci * ntags + t];
          if (skip_self && nn == mloc) continue;
          const double* rn = &s->pos[3 * (size_t)(first + nn)];
          double dx = rn[0] - rm[0];
          double dy = rn[1] - rm[1];
          double dz = rn[2] - rm[2];
          dx -= side * floor(side_inv * dx + 0.5);
          dy -= side * floor(side_inv * dy + 0.5);
          dz -= side * floor(side_inv * dz + 0.5);
          double dist2 = dx * dx + dy * dy + dz * dz;
          if (dist2 < rc2) {
            double dist_inv = sqrt(1.0 / dist2);
            double nf;
            orc_potential(&s->p, sigma, dist_inv, pot, &nf);
            double px = nf * dist_inv * dx;
            double py = nf * dist_inv * dy;
            double pz = nf * dist_inv * dz;
            *fx += px;
            *fy += py;
            *fz += pz;
            *vir -= 0.5 * (dx * px + dy * py + dz * pz);
            if (nbr) {
              if (*nn_count < maxn) nbr[*nn_count] = first + nn;
              (*nn_count)++;
            }
          }
        }
      }
}

static void forces_one(const orc_sys* s, int32_t m, double* out3, double* vir, double* pot,
                       int32_t* nbr, int32_t* nn_count, int32_t maxn) {
  int big = m < s->nb;
  double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0, u = 0.0;
  double sbs = 0.5 * (s->p.sigma_big + s->p.sigma_sml);
  if (big) {
    scan_grid(s, m, 0, s->g.nrange_bb, s->p.sigma_big, s->g.rc_bb * s->g.rc_bb, 1, &fx, &fy,
              &fz, &v, &u, nbr, nn_count, maxn);
    scan_grid(s, m, 1, s->g.nrange_bs, sbs, s->g.rc_bs * s->g.rc_bs, 0, &fx, &fy, &fz, &v, &u,
              nbr, nn_count, maxn);
  } else {
    scan_grid(s, m, 1, s->g.nrange_ss, s->p.sigma_sml, s->g.rc_ss * s->g.rc_ss, 1, &fx, &fy,
              &fz, &v, &u, nbr, nn_count, maxn);
    scan_grid(s, m, 0, s->g.nrange_sb, sbs, s->g.rc_bs * s->g.rc_bs, 0, &fx, &fy, &fz, &v, &u,
              nbr, nn_count, maxn);
  }
  out3[0] = fx;
  out3[1] = fy;
  out3[2] = fz;
  *vir = v;
  *pot = u;
}

int orc_compute_forces(orc_sys* s) {
  for (int32_t m = 0; m < s->n; m++)
    forces_one(s, m, &s->f[3 * (size_t)m], &s->vir[m], &s->pot[m], NULL, NULL, 0);
  return BD_OK;
}

int orc_neighbours(orc_sys* s, int32_t maxn, int32_t* counts, int32_t* list) {
  int rc = BD_OK;
  for (int32_t m = 0; m < s->n; m++) {
    double f3[3], v, u;
    int32_t c = 0;
    forces_one(s, m, f3, &v, &u, &list[(size_t)m * maxn], &c, maxn);
    counts[m] = c;
    if (c > maxn) rc = BD_E_ARG;
  }
  return rc;
}

int orc_forces_brute(const bd_params* p, const double* xyz, double* f, double* vir,
                     double* pot) {
  bd_geometry g;
  int rc = orc_geometry(p, &g);
  if (rc != BD_OK) return rc;
  int32_t n = p->n_big + p->n_sml;
  double side = p->side, side_inv = 1.0 / p->side;
  double sbs = 0.5 * (p->sigma_big + p->sigma_sml);
  for (int32_t m = 0; m < n; m++) {
    double fx = 0, fy = 0, fz = 0, v = 0, u = 0;
    for (int32_t k = 0; k < n; k++) {
      if (k == m) continue;
      int bm = m < p->n_big, bk = k < p->n_big;
      double sigma = bm && bk ? p->sigma_big : (!bm && !bk ? p->sigma_sml : sbs);
      double rc = bm && bk ? g.rc_bb : (!bm && !bk ? g.rc_ss : g.rc_bs);
      double dx = xyz[3 * k] - xyz[3 * m];
      double dy = xyz[3 * k + 1] - xyz[3 * m + 1];
      double dz = xyz[3 * k + 2] - xyz[3 * m + 2];
      dx -= side * floor(side_inv * dx + 0.5);
      dy -= side * floor(side_inv * dy + 0.5);
      dz -= side * floor(side_inv * dz + 0.5);
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < rc * rc) {
        double di = sqrt(1.0 / d2), nf;
        orc_potential(p, sigma, di, &u, &nf);
        double px = nf * di * dx, py = nf * di * dy, pz = nf * di * dz;
        fx += px;
        fy += py;
        fz += pz;
        v -= 0.5 * (dx * px + dy * py + dz * pz);
      }
    }
    f[3 * m] = fx;
    f[3 * m + 1] = fy;
    f[3 * m + 2] = fz;
    vir[m] = v;
    pot[m] = u;
  }
  return BD_OK;
}

/* ----------------------------------------------------------------------
 * update_ermak_hst (update_ermak.cu:8-77) with rannew64 lanes in place of
 * gsl_ran_gaussian.  Particle p (big first) is served by lane p % L; every
 * step each lane takes 4*ceil(N/L) draws, and the k-th particle of the lane
 * uses draws 4k..4k+3 (draws of absent particles are discarded), so all
 * lanes stay in lockstep.
 * -------------------------------------------------------------------- */
int orc_update(orc_sys* s) {
  double T = s->temp;
  double side = s->p.side;
  double side_inv = 1.0 / side;
  double dt = s->p.dt;
  int bad = 0;
  for (int32_t l = 0; l < s->lanes; l++) {
    for (int32_t k = 0; k < s->per_lane; k++) {
      uint64_t d[4];
      for (int j = 0; j < 4; j++) d[j] = orc_rn64_u64(&s->rng[l]);
      int64_t pidx = (int64_t)l + (int64_t)k * s->lanes;
      if (pidx >= s->n) continue;
      size_t p = (size_t)pidx;
      double diameter = pidx < s->nb ? s->p.sigma_big : s->p.sigma_sml;
      double dinv = 1.0 / diameter;
      double amplitude = sqrt(2.0 * T * dinv * dt);
      double xi[3];
      orc_gauss3(d, xi);
      for (int a = 0; a < 3; a++) {
        double drr = s->f[3 * p + a] * dinv * dt + amplitude * xi[a];
        double r = s->pos[3 * p + a] + drr;
        s->raw[3 * p + a] = s->raw[3 * p + a] + drr;
        r -= side * floor(side_inv * r);
        s->pos[3 * p + a] = r;
        if (!isfinite(r)) bad = 1;
      }
    }
  }
  s->ni++;
  s->temp = orc_temperature(&s->p, s->ni);
  return bad ? BD_E_NAN : BD_OK;
}

/* main.cu:494-586: update -> temperature -> cell locate -> forces */
int orc_step(orc_sys* s, int32_t nsteps) {
  for (int32_t i = 0; i < nsteps; i++) {
    int rc = orc_update(s);
    if (rc != BD_OK) return rc;
    rc = orc_cell_locate(s);
    if (rc != BD_OK) return rc;
    orc_compute_forces(s);
  }
  return BD_OK;
}

int orc_get_rng_lane(const orc_sys* s, int32_t lane, uint64_t* ring250, uint64_t* lcg,
                     int32_t* np, int32_t* nq) {
  if (lane < 0 || lane >= s->lanes) return BD_E_ARG;
  memcpy(ring250, s->rng[lane].ring, sizeof(uint64_t) * ORC_NP);
  *lcg = s->rng[lane].lcg;
  *np = s->rng[lane].np;
  *nq = s->rng[lane].nq;
  return BD_OK;
}

/* ----------------------------------------------------------------------
 * get_gder_hst (get_gder.cu:7-63) for (b,b), (b,s), (s,b), (s,s), counted
 * as integers (the reference adds 1.0f, exact below 2^24).
 * -------------------------------------------------------------------- */
int orc_gder(const orc_sys* s, int32_t nbins, double range, uint64_t* counts) {
  if (nbins < 1 || !(range > 0.0)) return BD_E_ARG;
  double side = s->p.side, side_inv = 1.0 / side;
  double bin_size = range / (double)nbins;
  for (int pr = 0; pr < 4; pr++) {
    int ta = pr >> 1, tb = pr & 1; /* 0 = big, 1 = sml */
    int32_t fa = ta ? s->nb : 0, na = ta ? s->ns : s->nb;
    int32_t fb = tb ? s->nb : 0, nbn = tb ? s->ns : s->nb;
    uint64_t* h = &counts[(size_t)pr * nbins];
    for (int32_t m = 0; m < na; m++) {
      const double* rm = &s->pos[3 * (size_t)(fa + m)];
      for (int32_t nn = 0; nn < nbn; nn++) {
        if (ta == tb && m == nn) continue;
        const double* rn = &s->pos[3 * (size_t)(fb + nn)];
        double dx = rm[0] - rn[0];
        double dy = rm[1] - rn[1];
        double dz = rm[2] - rn[2];
        dx -= side * floor(side_inv * dx + 0.5);
        dy -= side * floor(side_inv * dy + 0.5);
        dz -= side * floor(side_inv * dz + 0.5);
        double dist = sqrt(dx * dx + dy * dy + dz * dz);
        if (dist < range) {
          int32_t bin = (int32_t)(dist / bin_size);
          if (bin == nbins) bin--;
          h[bin]++;
        }
      }
    }
  }
  return BD_OK;
}

/* main.cu:849-877 for one sample (before the division by n_configs). */
void orc_gder_normalize(const bd_params* p, int32_t nbins, double range, const uint64_t* counts,
                        double* g) {
  double bs = range / (double)nbins;
  double volume = p->side * p->side * p->side;
  double xb = (double)p->n_big, xs = (double)p->n_sml;
  for (int32_t nb = 0; nb < nbins; nb++) {
    double xnb = (double)nb;
    double shell = 4.0 * PI_REF * bs * bs * bs * ((1.0 / 3.0) + xnb * xnb + xnb);
    double sigma = p->sigma_big;
    double vfree = volume - (4.0 / 3.0) * PI_REF * sigma * sigma * sigma;
    g[nb] = (double)counts[nb] / shell / (xb * (xb - 1.0) / vfree);
    sigma = 0.5 * (p->sigma_big + p->sigma_sml);
    vfree = volume - (4.0 / 3.0) * PI_REF * sigma * sigma * sigma;
    g[nbins + nb] = (double)counts[nbins + nb] / shell / (xb * xs / vfree);
    g[2 * nbins + nb] = (double)counts[2 * nbins + nb] / shell / (xs * xb / vfree);
    sigma = p->sigma_sml;
    vfree = volume - (4.0 / 3.0) * PI_REF * sigma * sigma * sigma;
    g[3 * nbins + nb] = (double)counts[3 * nbins + nb] / shell / (xs * (xs - 1.0) / vfree);
  }
}

/* ----------------------------------------------------------------------
 * fself_isotropic (fself.cu:3-21), MSD (main.cu:776-797), energies
 * (main.cu:627-638) and Z (main.cu:801-806).
 * Deviation: sin(x)/x at |dr| == 0 is taken as its limit 1 (the reference
 * divides 0 by 0).
 * -------------------------------------------------------------------- */
int orc_observables(const orc_sys* s, double q, bd_obs* o) {
  double fs = 0.0, mb = 0.0, ms = 0.0, eb = 0.0, es = 0.0, vir = 0.0;
  for (int32_t m = 0; m < s->n; m++) {
    double dx = s->raw[3 * (size_t)m] - s->ini[3 * (size_t)m];
    double dy = s->raw[3 * (size_t)m + 1] - s->ini[3 * (size_t)m + 1];
    double dz = s->raw[3 * (size_t)m + 2] - s->ini[3 * (size_t)m + 2];
    double d2 = dx * dx + dy * dy + dz * dz;
    if (m < s->nb) {
      double mag = sqrt(d2);
      double a = q * mag;
      fs += a == 0.0 ? 1.0 : sin(a) / a;
      mb += d2;
      eb += 0.5 * s->pot[m];
    } else {
      ms += d2;
      es += 0.5 * s->pot[m];
    }
    vir += s->vir[m];
  }
  o->fself_big = fs / (double)s->nb;
  o->msd_big = mb / (double)s->nb;
  o->msd_sml = ms / (double)s->ns;
  o->epot_big = eb / (double)s->nb;
  o->epot_sml = es / (double)s->ns;
  o->virial = vir;
  o->temperature = s->temp;
  o->z = 1.0 + vir / (3.0 * (double)s->n * s->temp);
  return BD_OK;
}

/* ======================================================================
 * Force contraction, Contracting_Vector3D.cu:57-164 (add_statistics) in
 * FP64.  ata is 3 x bins x bins (x, y, z), atb is 3 x bins; both are added
 * to.  side <= 0 means no periodic images (the std::nullopt case).
 * Deviation: a neighbour whose bin rounds to `bins` goes to bins-1 (the
 * reference would write past the end).
 * ==================================================================== */
int orc_cv3d_accumulate(int32_t bins, double range, const double* xyz, const double* f,
                        int32_t n, double side, double* ata, double* atb) {
  if (bins < 1 || !(range > 0.0) || n < 1) return BD_E_ARG;
  double bin_size = range / (double)bins;
  double bin_size_inv = 1.0 / bin_size;
  double range_sq = range * range;
  double side_inv = side > 0.0 ? 1.0 / side : 0.0;
  double* hit = (double*)calloc((size_t)bins * 3, sizeof(double));
  size_t bb = (size_t)bins * (size_t)bins;
  for (int32_t m = 0; m < n; m++) {
    int neighs = 0;
    for (int32_t nn = 0; nn < n; nn++) {
      if (nn == m) continue;
      double dx = xyz[3 * m] - xyz[3 * nn];
      double dy = xyz[3 * m + 1] - xyz[3 * nn + 1];
      double dz = xyz[3 * m + 2] - xyz[3 * nn + 2];
      if (side > 0.0) {
        dx -= side * floor(side_inv * dx + 0.5);
        dy -= side * floor(side_inv * dy + 0.5);
        dz -= side * floor(side_inv * dz + 0.5);
      }
      double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < range_sq) {
        double dist = sqrt(d2);
        int32_t nb = (int32_t)(bin_size_inv * dist);
        if (nb >= bins) nb = bins - 1;
        hit[3 * nb] += dx / dist;
        hit[3 * nb + 1] += dy / dist;
        hit[3 * nb + 2] += dz / dist;
        neighs++;
      }
    }
    if (neighs == 0) continue;
    for (int a = 0; a < 3; a++) {
      double* A = &ata[(size_t)a * bb];
      for (int32_t i = 0; i < bins; i++) {
        double hi = hit[3 * i + a];
        if (hi == 0.0) continue; /* a zero row adds +-0: no change */
        for (int32_t j = 0; j <= i; j++) {
          double aux = hi * hit[3 * j + a];
          A[(size_t)i * bins + j] += aux;
          if (i != j) A[(size_t)j * bins + i] += aux;
        }
      }
      for (int32_t i = 0; i < bins; i++) atb[(size_t)a * bins + i] += hit[3 * i + a] * f[3 * m + a];
    }
    memset(hit, 0, sizeof(double) * 3 * (size_t)bins);
  }
  free(hit);
  return BD_OK;
}

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (Golub &
 * Van Loan, "Matrix Computations", sec. 8.5): a = V diag(w) V^T. */
int orc_jacobi_eig(int32_t n, const double* a_in, double* w, double* v) {
  size_t nn = (size_t)n * (size_t)n;
  double* a = (double*)malloc(sizeof(double) * nn);
  memcpy(a, a_in, sizeof(double) * nn);
  for (size_t i = 0; i < nn; i++) v[i] = 0.0;
  for (int32_t i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, diag = 0.0;
    for (int32_t i = 0; i < n; i++) {
      diag += a[(size_t)i * n + i] * a[(size_t)i * n + i];
      for (int32_t j = i + 1; j < n; j++) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int32_t p = 0; p < n - 1; p++)
      for (int32_t q = p + 1; q < n; q++) {
        double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0);
        double sn = t * c;
        for (int32_t k = 0; k < n; k++) {
          double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - sn * akq;
          a[(size_t)k * n + q] = sn * akp + c * akq;
        }
        for (int32_t k = 0; k < n; k++) {
          double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - sn * aqk;
          a[(size_t)q * n + k] = sn * apk + c * aqk;
        }
        for (int32_t k = 0; k < n; k++) {
          double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - sn * vkq;
          v[(size_t)k * n + q] = sn * vkp + c * vkq;
        }
      }
  }
  for (int32_t i = 0; i < n; i++) w[i] = a[(size_t)i * n + i];
  free(a);
  return BD_OK;
}

/* calculate_vector (Contracting_Vector3D.cu:166-297): per axis the
 * truncated pseudo-inverse solution x = V S^+ U^T b with singular values
 * kept when |w/w_max| > 1e-6, then the average over the three axes.  For a
 * symmetric matrix S = |eigenvalues| and the solution equals
 * sum_k q_k (q_k . b) / lambda_k over the kept eigenpairs. */
int orc_cv3d_solve(int32_t bins, const double* ata, const double* atb, double* out) {
  size_t bb = (size_t)bins * (size_t)bins;
  double* w = (double*)malloc(sizeof(double) * (size_t)bins);
  double* v = (double*)malloc(sizeof(double) * bb);
  for (int32_t i = 0; i < bins; i++) out[i] = 0.0;
  for (int a = 0; a < 3; a++) {
    orc_jacobi_eig(bins, &ata[(size_t)a * bb], w, v);
    double wmax = 0.0;
    for (int32_t i = 0; i < bins; i++)
      if (fabs(w[i]) > wmax) wmax = fabs(w[i]);
    const double* b = &atb[(size_t)a * bins];
    for (int32_t k = 0; k < bins; k++) {
      if (!(wmax > 0.0) || !(fabs(w[k] / wmax) > 1.0e-6)) continue;
      double proj = 0.0;
      for (int32_t i = 0; i < bins; i++) proj += v[(size_t)i * bins + k] * b[i];
      double c = proj / w[k];
      for (int32_t i = 0; i < bins; i++) out[i] += v[(size_t)i * bins + k] * c;
    }
  }
  for (int32_t i = 0; i < bins; i++) out[i] /= 3.0;
  free(w);
  free(v);
  return BD_OK;
}
