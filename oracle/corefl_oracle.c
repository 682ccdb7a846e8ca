/*
 * corefl_oracle.c -- CPU restatement of COREFL's per-stage RHS + SSP-RK3 update.
 *
 * TEST INFRASTRUCTURE ONLY (see corefl_oracle.h).  Scalar C, one point at a
 * time, in the reference's operation order so that the HIP product can be
 * compared against it at 1e-11 normwise relative per variable.
 *
 * Transcendentals (pow, exp, log, sqrt) come from the host libm; the
 * reference's CUDA norm3d/rnorm3d are written as sqrt(x*x+y*y+z*z) and its
 * reciprocal, from their definition.
 */
#include "corefl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* physical constants, src/Constants.h:13-35 */
static const double R_u = 8314.462618;
static const double p_atm = 101325;
static const double gamma_air = 1.4;
static const double mw_air = 28.9647;
static const double pi_c = 3.141592653589793;
static const double eps_weno = 1e-40; /* src/WENO.cu:8 */

#define R_AIR (R_u / mw_air)

/* SSP-RK3 coefficients, src/RK.cuh:12-16 */
static const double rk_a[3] = {1.0, 0.75, 1.0 / 3.0};
static const double rk_b[3] = {0.0, 0.25, 2.0 / 3.0};
static const double rk_c[3] = {1.0, 0.25, 2.0 / 3.0};

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double norm3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

/* ------------------------------------------------------------------------ */
/* indexing                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  long sx, sxy, sz, off;
} lay_t;

static lay_t layout(const oc_zone *z) {
  lay_t L;
  const long g = z->ngg;
  L.sx = z->mx + 2 * g;
  L.sxy = L.sx * (z->my + 2 * g);
  L.sz = L.sxy * (z->mz + 2 * g);
  L.off = g * (L.sxy + L.sx + 1);
  return L;
}
#define IX(L, i, j, k) ((long)(k) * (L).sxy + (long)(j) * (L).sx + (long)(i) + (L).off)
#define DQX(z, i, j, k) (((long)(k) * (z)->my + (j)) * (long)(z)->mx + (i))
#define DQSZ(z) ((long)(z)->mx * (z)->my * (z)->mz)

/* ------------------------------------------------------------------------ */
/* thermodynamics: NASA-7, Combustion2Part build, src/Thermo.cu:386-631      */
/* ------------------------------------------------------------------------ */
static double nasa7_h_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
  return c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
}
static double nasa7_cp_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
  return c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
}

/* NASA-9 (src/Thermo.cu:426-456, 501-521, 573-596, 628-653): range j of species
 * i is the first with t_cut[j] <= t <= t_cut[j+1]; outside [t_cut[0],
 * t_cut[n]] enthalpy is extrapolated linearly with cp at the bound, cp is held
 * and G/RT is evaluated at the bound. */
static double nasa9_h_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt, itt = 1.0 / tt;
  return -c[0] * itt + c[1] * log(tt) + c[2] * tt + 0.5 * c[3] * tt2 + c[4] * tt3 / 3 + 0.25 * c[5] * tt4 +
         0.2 * c[6] * tt5 + c[7];
}
static double nasa9_cp_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, itt2 = 1.0 / tt2, itt = 1.0 / tt;
  return c[0] * itt2 + c[1] * itt + c[2] + c[3] * tt + c[4] * tt2 + c[5] * tt3 + c[6] * tt4;
}
static double nasa9_g_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, itt = 1 / tt, lntt = log(tt), itt2 = itt * itt;
  return -0.5 * c[0] * itt2 + c[1] * itt * (lntt + 1) + c[2] * (1.0 - lntt) - 0.5 * c[3] * tt - c[4] * tt2 / 6.0 -
         c[5] * tt3 / 12.0 - c[6] * tt4 * 0.05 + c[7] * itt - c[8];
}
/* -1 below the first cut, n above the last, else the range index */
static int nasa9_range(const oc_tables *tb, int i, double t) {
  const int n = tb->n_range[i];
  if (t < tb->t_cut[i][0]) return -1;
  if (t > tb->t_cut[i][n]) return n;
  for (int j = 0; j < n; ++j)
    if (tb->t_cut[i][j] <= t && t <= tb->t_cut[i][j + 1]) return j;
  return n;
}
static void nasa9_h_cp(const oc_tables *tb, int i, double t, double *h, double *cp) {
  const int n = tb->n_range[i], j = nasa9_range(tb, i, t);
  if (j < 0 || j >= n) {
    const int jj = j < 0 ? 0 : n - 1;
    const double tt = j < 0 ? tb->t_cut[i][0] : tb->t_cut[i][n];
    const double c = nasa9_cp_at(tb->poly9[i][jj], tt);
    *h = nasa9_h_at(tb->poly9[i][jj], tt) + c * (t - tt);
    *cp = c;
  } else {
    const double *c = tb->poly9[i][j];
    const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t, it = 1.0 / t, lnT = log(t), it2 = it * it;
    *h = -c[0] * it + c[1] * lnT + c[2] * t + 0.5 * c[3] * t2 + c[4] * t3 / 3 + 0.25 * c[5] * t4 + 0.2 * c[6] * t5 +
         c[7];
    *cp = c[0] * it2 + c[1] * it + c[2] + c[3] * t + c[4] * t2 + c[5] * t3 + c[6] * t4;
  }
}

/* compute_enthalpy, src/Thermo.cu:386-423 (NASA-7 branch) */
void oc_enthalpy(const oc_tables *tb, double t, double *h) {
  if (tb->nasa7or9 == 9) {
    for (int i = 0; i < tb->n_spec; ++i) {
      double cp;
      nasa9_h_cp(tb, i, t, &h[i], &cp);
      h[i] *= tb->gas_const[i];
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < tb->n_spec; ++i) {
    if (t < tb->t_low[i]) {
      const double tt = tb->t_low[i];
      const double *c = tb->low_coeff[i];
      h[i] = nasa7_h_at(c, tt);
      const double cp = nasa7_cp_at(c, tt);
      h[i] += cp * (t - tt);
    } else if (t > tb->t_high[i]) {
      const double tt = tb->t_high[i];
      const double *c = tb->high_coeff[i];
      h[i] = nasa7_h_at(c, tt);
      const double cp = nasa7_cp_at(c, tt);
      h[i] += cp * (t - tt);
    } else {
      const double *c = t < tb->t_mid[i] ? tb->low_coeff[i] : tb->high_coeff[i];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
    }
    h[i] *= tb->gas_const[i];
  }
}

/* compute_enthalpy_and_cp, src/Thermo.cu:465-500 (NASA-7 branch) */
void oc_enthalpy_and_cp(const oc_tables *tb, double t, double *h, double *cp) {
  if (tb->nasa7or9 == 9) {
    for (int i = 0; i < tb->n_spec; ++i) {
      nasa9_h_cp(tb, i, t, &h[i], &cp[i]);
      h[i] *= tb->gas_const[i];
      cp[i] *= tb->gas_const[i];
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < tb->n_spec; ++i) {
    if (t < tb->t_low[i]) {
      const double tt = tb->t_low[i];
      const double *c = tb->low_coeff[i];
      h[i] = nasa7_h_at(c, tt);
      cp[i] = nasa7_cp_at(c, tt);
      h[i] += cp[i] * (t - tt);
    } else if (t > tb->t_high[i]) {
      const double tt = tb->t_high[i];
      const double *c = tb->high_coeff[i];
      h[i] = nasa7_h_at(c, tt);
      cp[i] = nasa7_cp_at(c, tt);
      h[i] += cp[i] * (t - tt);
    } else {
      const double *c = t < tb->t_mid[i] ? tb->low_coeff[i] : tb->high_coeff[i];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= tb->gas_const[i];
    cp[i] *= tb->gas_const[i];
  }
}

/* compute_cp, src/Thermo.cu:551-572 (NASA-7 branch) */
void oc_cp(const oc_tables *tb, double t, double *cp) {
  if (tb->nasa7or9 == 9) {
    for (int i = 0; i < tb->n_spec; ++i) {
      double h;
      nasa9_h_cp(tb, i, t, &h, &cp[i]);
      cp[i] *= tb->gas_const[i];
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
  for (int i = 0; i < tb->n_spec; ++i) {
    if (t < tb->t_low[i]) {
      cp[i] = nasa7_cp_at(tb->low_coeff[i], tb->t_low[i]);
    } else if (t > tb->t_high[i]) {
      cp[i] = nasa7_cp_at(tb->high_coeff[i], tb->t_high[i]);
    } else {
      const double *c = t < tb->t_mid[i] ? tb->low_coeff[i] : tb->high_coeff[i];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= tb->gas_const[i];
  }
}

/* compute_gibbs_div_rt, src/Thermo.cu:600-626 (NASA-7 branch) */
void oc_gibbs_div_rt(const oc_tables *tb, double t, double *g) {
  if (tb->nasa7or9 == 9) {
    for (int i = 0; i < tb->n_spec; ++i) {
      const int n = tb->n_range[i], j = nasa9_range(tb, i, t);
      if (j < 0) g[i] = nasa9_g_at(tb->poly9[i][0], tb->t_cut[i][0]);
      else if (j >= n) g[i] = nasa9_g_at(tb->poly9[i][n - 1], tb->t_cut[i][n]);
      else g[i] = nasa9_g_at(tb->poly9[i][j], t);
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
  for (int i = 0; i < tb->n_spec; ++i) {
    if (t < tb->t_low[i] || t > tb->t_high[i]) {
      const double tt = t < tb->t_low[i] ? tb->t_low[i] : tb->t_high[i];
      const double *c = t < tb->t_low[i] ? tb->low_coeff[i] : tb->high_coeff[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
      g[i] = c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
             c[5] * tt_inv - c[6];
    } else {
      const double *c = t < tb->t_mid[i] ? tb->low_coeff[i] : tb->high_coeff[i];
      g[i] = c[0] * (1.0 - log_t) - 0.5 * c[1] * t - c[2] * t2 / 6.0 - c[3] * t3 / 12.0 - c[4] * t4 * 0.05 +
             c[5] * t_inv - c[6];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* two-temperature energy modes and V-T relaxation, src/Thermo.cu:635-901    */
/* (per unit mass; the vibrational-electronic temperature is clamped >= 1 K) */
/* ------------------------------------------------------------------------ */
static double vib_energy(const oc_tables *tb, int i, double t) {
  const double temp = dmax(t, 1.0);
  double e = 0.0;
  if (tb->theta_v[i] > 0) {
    const double x = tb->theta_v[i] / temp;
    if (x < 200) e += tb->gas_const[i] * tb->theta_v[i] / expm1(x);
  }
  return e;
}
static double vib_cv(const oc_tables *tb, int i, double t) {
  const double temp = dmax(t, 1.0);
  double c = 0.0;
  if (tb->theta_v[i] > 0) {
    const double x = tb->theta_v[i] / temp;
    if (x < 200) {
      const double ex = exp(x), d = ex - 1.0;
      c += tb->gas_const[i] * x * x * ex / (d * d);
    }
  }
  return c;
}
double oc_ve_energy(const oc_tables *tb, int i, double t) {
  if (!tb->two_temperature) return 0.0;
  const double temp = dmax(t, 1.0);
  double e = vib_energy(tb, i, temp);
  const int n = tb->n_elec[i];
  if (n > 0) {
    double z = 0.0, w = 0.0;
    for (int j = 0; j < n; ++j) {
      const double th = tb->elec_theta[i][j], boltz = tb->elec_g[i][j] * exp(-th / temp);
      z += boltz;
      w += boltz * th;
    }
    if (z > 0) e += tb->gas_const[i] * w / z;
  }
  return e;
}
double oc_ve_cv(const oc_tables *tb, int i, double t) {
  if (!tb->two_temperature) return 0.0;
  const double temp = dmax(t, 1.0);
  double c = vib_cv(tb, i, temp);
  const int n = tb->n_elec[i];
  if (n > 0) {
    double z = 0.0, b = 0.0, cc = 0.0;
    for (int j = 0; j < n; ++j) {
      const double th = tb->elec_theta[i][j], boltz = tb->elec_g[i][j] * exp(-th / temp);
      z += boltz;
      b += boltz * th;
      cc += boltz * th * th;
    }
    if (z > 0) {
      const double mean = b / z;
      c += tb->gas_const[i] * (cc / z - mean * mean) / (temp * temp);
    }
  }
  return c;
}
/* compute_mixture_ve_energy, src/Thermo.cu:716-729 */
double oc_mixture_ve_energy(const oc_tables *tb, double t, const double *y, double *cv_ve) {
  double e = 0.0, c = 0.0;
  if (tb->two_temperature) {
    for (int i = 0; i < tb->n_spec; ++i) {
      e += y[i] * oc_ve_energy(tb, i, t);
      c += y[i] * oc_ve_cv(tb, i, t);
    }
  }
  if (cv_ve) *cv_ve = c;
  return e;
}
/* compute_mixture_vib_energy, src/Thermo.cu:731-744 */
double oc_mixture_vib_energy(const oc_tables *tb, double t, const double *y, double *cv_v) {
  double e = 0.0, c = 0.0;
  if (tb->two_temperature) {
    for (int i = 0; i < tb->n_spec; ++i) {
      e += y[i] * vib_energy(tb, i, t);
      c += y[i] * vib_cv(tb, i, t);
    }
  }
  if (cv_v) *cv_v = c;
  return e;
}
/* compute_mixture_tr_energy, src/Thermo.cu:746-770: the equilibrium internal
 * energy less E_ve(T), with cv_tr, cp_tr and R of the mixture */
double oc_mixture_tr_energy(const oc_tables *tb, double t, const double *y, double *cv_tr, double *cp_tr,
                            double *r_mix) {
  double h[OC_MAX_SPEC], cp[OC_MAX_SPEC];
  oc_enthalpy_and_cp(tb, t, h, cp);
  double e = 0.0, cpm = 0.0, rm = 0.0;
  for (int i = 0; i < tb->n_spec; ++i) {
    cpm += y[i] * cp[i];
    rm += y[i] * tb->gas_const[i];
    e += y[i] * (h[i] - tb->gas_const[i] * t);
  }
  double cv_ve = 0.0;
  const double eve = oc_mixture_ve_energy(tb, t, y, &cv_ve);
  const double cv_total = cpm - rm;
  if (cv_tr) *cv_tr = cv_total - cv_ve;
  if (cp_tr) *cp_tr = cpm - cv_ve;
  if (r_mix) *r_mix = rm;
  return e - eve;
}
/* invert_tve_from_eve, src/Thermo.cu:772-787 (Newton, at most 80 steps) */
double oc_invert_tve(const oc_tables *tb, double eve_target, const double *y, double t_init) {
  double tve = dmax(t_init, 1.0);
  for (int it = 0; it < 80; ++it) {
    double cv_ve = 0.0;
    const double eve = oc_mixture_ve_energy(tb, tve, y, &cv_ve);
    const double res = eve - eve_target;
    if (fabs(res) < 1e-8 * dmax(1.0, fabs(eve_target))) break;
    const double next = dmax(1.0, tve - res / dmax(cv_ve, 1e-8));
    if (fabs(next - tve) < 1e-8 * dmax(1.0, tve)) {
      tve = next;
      break;
    }
    tve = next;
  }
  return tve;
}
/* Millikan-White pair time (with the default correlation when the pair has no
 * coefficients) and the species time with Park's high-temperature correction,
 * src/Thermo.cu:805-854 */
static double mw_pair_tau(const oc_tables *tb, int iv, int ip, double t, double p) {
  if (p <= 1e-16 || t <= 1.0) return 1e30;
  const double mv = 1.0 / dmax(tb->imw[iv], 1e-20), mp = 1.0 / dmax(tb->imw[ip], 1e-20);
  const double mu = mv * mp / dmax(mv + mp, 1e-20);
  double a = tb->lt_a[iv][ip], b = tb->lt_b[iv][ip];
  if (a <= 0.0 || b <= 0.0) {
    a = 1.16e-3 * sqrt(dmax(mu, 1e-20)) * pow(dmax(tb->theta_v[iv], 0.0), 4.0 / 3.0);
    b = 0.015 * pow(dmax(mu, 1e-20), 0.25);
  }
  const double ex = a * (pow(t, -1.0 / 3.0) - b) - 18.421;
  return dmax(exp(ex) * 101325.0 / p, 1e-30);
}
static double species_vt_tau(const oc_tables *tb, int iv, double rho, double t, double p, const double *y) {
  double conc = 0.0, den = 0.0;
  for (int j = 0; j < tb->n_spec; ++j) {
    const double yj = dmax(y[j], 0.0);
    if (yj <= 0.0) continue;
    const double mole = yj / (1.0 / dmax(tb->imw[j], 1e-20));
    conc += mole;
    den += mole / mw_pair_tau(tb, iv, j, t, p);
  }
  if (conc <= 0.0 || den <= 0.0) return 1e30;
  const double tau_mw = conc / den;
  const double sigma = dmax(tb->park_sigma[iv], 1e-30) * (2.5e9 / dmax(t * t, 1.0));
  const double n_total = rho * conc * 6.02214076e26;
  const double mv = 1.0 / dmax(tb->imw[iv], 1e-20);
  const double cs = sqrt(dmax(8.0 * 8314.462618 * t / (3.141592653589793 * mv), 1e-30));
  const double tau_park = 1.0 / dmax(sigma * cs * n_total, 1e-30);
  return dmax(tau_mw + tau_park, 1e-30);
}
/* compute_vt_relaxation_source, src/Thermo.cu:856-901 */
double oc_vt_source(const oc_tables *tb, double rho, double t, double tve, const double *y, double *eve_eq,
                    double *tau_eff) {
  if (!tb->two_temperature || rho <= 0.0) {
    if (eve_eq) *eve_eq = 0.0;
    if (tau_eff) *tau_eff = 1e30;
    return 0.0;
  }
  double rm = 0.0;
  for (int i = 0; i < tb->n_spec; ++i) rm += dmax(y[i], 0.0) * tb->gas_const[i];
  const double p = rho * rm * dmax(t, 1.0);
  double eq = 0.0, em = 0.0, src = 0.0;
  for (int i = 0; i < tb->n_spec; ++i) {
    if (tb->theta_v[i] <= 0.0) continue;
    const double yi = dmax(y[i], 0.0);
    if (yi <= 0.0) continue;
    const double e_eq = vib_energy(tb, i, t), e_i = vib_energy(tb, i, tve), d = e_eq - e_i;
    eq += yi * e_eq;
    em += yi * e_i;
    if (fabs(d) <= 1e-16) continue;
    const double tau = species_vt_tau(tb, i, rho, t, p, y);
    if (tau < 1e29) src += rho * yi * d / tau;
  }
  if (eve_eq) *eve_eq = eq;
  if (tau_eff) {
    const double dm = eq - em;
    *tau_eff = (fabs(dm) > 1e-16 && fabs(src) > 1e-16) ? dmax(rho * dm / src, 1e-30) : 1e30;
  }
  return src;
}

/* ------------------------------------------------------------------------ */
/* finite-rate chemistry, generic path, src/FiniteRateChem.cu:1505-1681      */
/* ------------------------------------------------------------------------ */
static double arrhenius(double t, double A, double b, double Ea) { return A * pow(t, b) * exp(-Ea / t); }

/* reaction (control) temperature, src/FiniteRateChem.cu:1490-1503: T^a Tve^b
 * through a smooth maximum with 800 K (width 80 K) when the model is on */
static double reaction_temperature(const oc_tables *tb, double t, double tve, double a, double b) {
  if (!tb->trx_model) return t;
  const double trx = pow(dmax(t, 1.0), a) * pow(dmax(tve, 1.0), b);
  return 0.5 * (trx + 800.0 + sqrt((trx - 800.0) * (trx - 800.0) + 80.0 * 80.0));
}

static void chem_rates(const oc_tables *tb, double rho, double t, double tve, const double *y, double *omega,
                       double *omega_d, double *q1, double *q2);

void oc_chem_source(const oc_tables *tb, double rho, double t, double tve, const double *y, double *omega,
                    double *omega_d) {
  double q1[OC_MAX_REAC], q2[OC_MAX_REAC];
  chem_rates(tb, rho, t, tve, y, omega, omega_d, q1, q2);
}

/* the source of oc_chem_source with the rates of progress q1 (forward) and
 * q2 (backward) of every reaction, mol/(cm^3 s) */
static void chem_rates(const oc_tables *tb, double rho, double t, double tve, const double *y, double *omega,
                       double *omega_d, double *q1, double *q2) {
  const int ns = tb->n_spec, nr = tb->n_reac;
  double c[OC_MAX_SPEC], kf[OC_MAX_REAC], kb[OC_MAX_REAC];
  /* concentration in mol/cm3, src/FiniteRateChem.cu:1419-1424 */
  for (int l = 0; l < ns; ++l) c[l] = rho * y[l] * tb->imw[l] * 1e-3;

  /* forward_reaction_rate, src/FiniteRateChem.cu:1505-1554 */
  for (int i = 0; i < nr; ++i) {
    const double thf = reaction_temperature(tb, t, tve, tb->tcf_a[i], tb->tcf_b[i]);
    const int type = tb->reac_type[i];
    kf[i] = arrhenius(thf, tb->A[i], tb->b[i], tb->Ea[i]);
    if (type == 3) {
      kf[i] += arrhenius(thf, tb->A2[i], tb->b2[i], tb->Ea2[i]);
    } else if (type > 3) {
      double cc = 0;
      for (int l = 0; l < ns; ++l) cc += c[l] * tb->third_body[i][l];
      if (type == 4) {
        kf[i] *= cc;
      } else {
        const double kf_low = arrhenius(thf, tb->A2[i], tb->b2[i], tb->Ea2[i]);
        const double kf_high = kf[i];
        if (kf_high < 1e-25 && kf_low < 1e-25) {
          kf[i] = 0;
          continue;
        }
        const double reduced_pressure = kf_low * cc / kf_high;
        double F = 1.0;
        if (type > 5) {
          double f_cent = (1 - tb->troe_alpha[i]) * exp(-thf / tb->troe_t3[i]) +
                          tb->troe_alpha[i] * exp(-thf / tb->troe_t1[i]);
          if (type == 7) f_cent += exp(-tb->troe_t2[i] / thf);
          const double logFc = log10(f_cent);
          const double cT = -0.4 - 0.67 * logFc;
          const double n = 0.75 - 1.27 * logFc;
          const double logPr = log10(reduced_pressure);
          const double tempo = (logPr + cT) / (n - 0.14 * (logPr + cT));
          const double p = logFc / (1.0 + tempo * tempo);
          F = pow(10, p);
        }
        kf[i] = kf_high * reduced_pressure / (1.0 + reduced_pressure) * F;
      }
    }
  }

  /* backward_reaction_rate with per-reaction temperatures, src/FiniteRateChem.cu:1560-1589 */
  if (tb->trx_model) {
    for (int i = 0; i < nr; ++i) {
      const double thb = reaction_temperature(tb, t, tve, tb->tcb_a[i], tb->tcb_b[i]);
      if (tb->reac_type[i] == 0) {
        kb[i] = 0;
        continue;
      }
      if (tb->rev_type[i] == 1) {
        kb[i] = arrhenius(thb, tb->A2[i], tb->b2[i], tb->Ea2[i]);
        if (tb->reac_type[i] == 4) {
          double cc = 0;
          for (int l = 0; l < ns; ++l) cc += c[l] * tb->third_body[i][l];
          kb[i] *= cc;
        }
        continue;
      }
      double g[OC_MAX_SPEC];
      oc_gibbs_div_rt(tb, thb, g);
      const double temp_t = p_atm / R_u * 1e-3 / thb;
      double d_gibbs = 0;
      for (int l = 0; l < ns; ++l) d_gibbs += g[l] * (tb->stoi_b[i][l] - tb->stoi_f[i][l]);
      const double kc = pow(temp_t, tb->order[i]) * exp(-d_gibbs);
      kb[i] = kf[i] / kc;
    }
  } else {
    /* backward_reaction_rate, src/FiniteRateChem.cu:1591-1636 (one temperature) */
    int n_gibbs = nr;
    for (int i = 0; i < nr; ++i) {
      if (tb->reac_type[i] == 0) {
        kb[i] = 0;
        --n_gibbs;
      } else if (tb->rev_type[i] == 1) {
        kb[i] = arrhenius(t, tb->A2[i], tb->b2[i], tb->Ea2[i]);
        if (tb->reac_type[i] == 4) {
          double cc = 0;
          for (int l = 0; l < ns; ++l) cc += c[l] * tb->third_body[i][l];
          kb[i] *= cc;
        }
        --n_gibbs;
      }
    }
    if (n_gibbs >= 1) {
      double g[OC_MAX_SPEC];
      oc_gibbs_div_rt(tb, t, g);
      const double temp_p = p_atm / R_u * 1e-3;
      const double temp_t = temp_p / t;
      for (int i = 0; i < nr; ++i) {
        if (tb->reac_type[i] != 2 && tb->reac_type[i] != 0) {
          double d_gibbs = 0;
          for (int l = 0; l < ns; ++l) d_gibbs += g[l] * (tb->stoi_b[i][l] - tb->stoi_f[i][l]);
          const double kc = pow(temp_t, tb->order[i]) * exp(-d_gibbs);
          kb[i] = kf[i] / kc;
        }
      }
    }
  }

  /* rate_of_progress, src/FiniteRateChem.cu:1638-1663 */
  for (int i = 0; i < nr; ++i) {
    q1[i] = 1.0;
    if (tb->reac_type[i] != 0) {
      q2[i] = 1.0;
      for (int j = 0; j < ns; ++j) {
        q1[i] *= pow(c[j], tb->stoi_f[i][j]);
        q2[i] *= pow(c[j], tb->stoi_b[i][j]);
      }
      q1[i] *= kf[i];
      q2[i] *= kb[i];
    } else {
      q2[i] = 0.0;
      for (int j = 0; j < ns; ++j) q1[i] *= pow(c[j], tb->stoi_f[i][j]);
      q1[i] *= kf[i];
    }
  }

  /* chemical_source, src/FiniteRateChem.cu:1665-1681 */
  for (int i = 0; i < ns; ++i) {
    double creation = 0;
    omega_d[i] = 0;
    for (int j = 0; j < nr; ++j) {
      creation += q2[j] * tb->stoi_f[j][i] + q1[j] * tb->stoi_b[j][i];
      omega_d[i] += q1[j] * tb->stoi_f[j][i] + q2[j] * tb->stoi_b[j][i];
    }
    creation *= 1e+3 / tb->imw[i];
    omega_d[i] *= 1e+3 / tb->imw[i];
    omega[i] = creation - omega_d[i];
  }
}

/* ------------------------------------------------------------------------ */
/* transport, src/Transport.cu:43-150, Omega_D src/Transport.cu:182-184       */
/* ------------------------------------------------------------------------ */
static double omega_d_fit(double t_red) { return 1.0548 * pow(t_red, -0.15504) + pow(t_red + 0.55909, -2.1705); }

void oc_transport(const oc_tables *tb, double temperature, double mw_total, const double *y, double *cp,
                  double *mul, double *kappa, double *rhoD) {
  const int n = tb->n_spec;
  double mw[OC_MAX_SPEC], X[OC_MAX_SPEC], vis[OC_MAX_SPEC], num[OC_MAX_SPEC], den[OC_MAX_SPEC];
  for (int m = 0; m < n; ++m) mw[m] = 1.0 / tb->imw[m];
  const double sqrt_t = sqrt(temperature);
  for (int l = 0; l < n; ++l) {
    X[l] = y[l] * mw_total / mw[l];
    const double t_dl = temperature * tb->LJ_potent_inv[l];
    vis[l] = tb->vis_coeff[l] * sqrt_t / (1.147 * pow(t_dl, -0.145) + pow(t_dl + 0.5, -2));
  }
  const double t_3over2 = temperature * sqrt_t;
  double viscosity = 0, conductivity = 0;
  for (int m = 0; m < n; ++m) {
    double R = tb->gas_const[m];
    double lambda = 15 * 0.25 * vis[m] * R;
    const int geom = tb->geometry[m];
    if (geom == 1 || geom == 2) {
      const double t_red = temperature * tb->kb_over_eps_jk[m][m];
      const double tRedInv = 1 / t_red;
      const double tmp = sqrt(tRedInv * pi_c) * pi_c;
      const double FT = 1 + 0.5 * tmp + (2 + 0.25 * pi_c * pi_c) * tRedInv + tmp * tRedInv;
      const double rD = tb->binary_diffusivity_coeff[m][m] * t_3over2 / (omega_d_fit(t_red) * R * temperature);
      if (geom == 1) {
        const double ADivPiB = (2.5 * vis[m] - rD) /
                               ((pi_c * tb->ZRotF298[m] / FT * vis[m] + 10.0 / 3 * vis[m] + 2.0 * rD));
        lambda += -5 * ADivPiB * vis[m] * R + rD * cp[m] + rD * (2 * ADivPiB - 2.5) * R;
      } else {
        const double ADivPiB = (2.5 * vis[m] - rD) /
                               ((pi_c * tb->ZRotF298[m] / FT * vis[m] + 5 * vis[m] + 2.0 * rD));
        lambda += -7.5 * ADivPiB * vis[m] * R + rD * cp[m] + rD * (3 * ADivPiB - 2.5) * R;
      }
    }
    double sum = 0;
    R = sqrt(vis[m]);
    for (int k = 0; k < n; ++k) {
      double pf = 1.0;
      if (m != k) {
        const double numer = 1.0 + (1.0 / sqrt(vis[k])) * (R * tb->WjDivWi_to_One4th[m][k]);
        pf = numer * numer * tb->sqrt_WiDivWjPl1Mul8[m][k];
      }
      sum += pf * X[k];
    }
    viscosity += vis[m] * X[m] / sum;
    conductivity += lambda * X[m] / (1.065 * sum - 0.065 * X[m]);
  }
  *mul = viscosity;
  *kappa = conductivity;

  const double eps = 1e-12;
  double s = 0;
  for (int l = 0; l < n; ++l) {
    const double tmp = (X[l] + eps) * mw[l];
    num[l] = tmp;
    s += tmp;
    den[l] = 0;
  }
  for (int l = 0; l < n; ++l) num[l] = s - num[l];
  for (int l = 0; l < n; ++l) {
    for (int k = l + 1; k < n; ++k) {
      const double tmp = omega_d_fit(temperature * tb->kb_over_eps_jk[l][k]) /
                         (tb->binary_diffusivity_coeff[l][k] * t_3over2);
      den[l] += (X[k] + eps) * tmp;
      den[k] += (X[l] + eps) * tmp;
    }
  }
  const double inv = 1.0 / (temperature * R_u);
  for (int l = 0; l < n; ++l) rhoD[l] = num[l] * inv / den[l];
}

/* Sutherland law, src/Transport.cu:6-9 */
static double sutherland(double t) { return 1.7894e-5 * pow(t / 288.16, 1.5) * (288.16 + 110) / (t + 110); }

/* ------------------------------------------------------------------------ */
/* field operations                                                          */
/* ------------------------------------------------------------------------ */

/* compute_cv_from_bv_1_point + compute_total_energy (1T),
 * src/FieldOperation.cuh:90-161 */
/* sync_two_temperature_state_from_sv, src/FieldOperation.cuh:28-57: Tve from
 * the E_ve scalar (or E_ve from Tve when the scalar is unset or unusable) */
static void sync_two_temperature(oc_zone *z, const lay_t *L, const oc_tables *tb, long p) {
  double y[OC_MAX_SPEC];
  for (int l = 0; l < tb->n_spec; ++l) y[l] = z->sv[p + l * L->sz];
  const double t = dmax(z->bv[p + 5 * L->sz], 1.0);
  double tve = z->tve[p];
  const double max_guess = dmax(1.0e7, 1000.0 * t);
  if (!isfinite(tve) || tve <= 0.0 || tve > max_guess) tve = t;
  double *eve = &z->sv[p + tb->n_spec * L->sz];
  if (!isfinite(*eve) || *eve <= 0.0) {
    *eve = oc_mixture_ve_energy(tb, tve, y, NULL);
  } else {
    const double back = oc_invert_tve(tb, *eve, y, tve);
    if (isfinite(back) && back > 0.0 && back <= max_guess) {
      tve = back;
    } else {
      tve = t;
      *eve = oc_mixture_ve_energy(tb, tve, y, NULL);
    }
  }
  z->tve[p] = tve;
}

/* compute_cv_from_bv_1_point + compute_total_energy, src/FieldOperation.cuh:90-180 */
static void cv_from_bv_point(oc_zone *z, const lay_t *L, const oc_run *r, const oc_tables *tb, long p) {
  const double rho = z->bv[p], u = z->bv[p + L->sz], v = z->bv[p + 2 * L->sz], w = z->bv[p + 3 * L->sz];
  z->cv[p] = rho;
  z->cv[p + L->sz] = rho * u;
  z->cv[p + 2 * L->sz] = rho * v;
  z->cv[p + 3 * L->sz] = rho * w;
  if (r->two_temperature) sync_two_temperature(z, L, tb, p);
  for (int l = 0; l < r->n_scalar; ++l) z->cv[p + (5 + l) * L->sz] = rho * z->sv[p + l * L->sz];
  const double V2 = u * u + v * v + w * w;
  double E = 0.5 * V2;
  if (r->mixture && r->two_temperature) {
    const double T = z->bv[p + 5 * L->sz];
    double h[OC_MAX_SPEC], cp[OC_MAX_SPEC], y[OC_MAX_SPEC];
    oc_enthalpy_and_cp(tb, T, h, cp);
    for (int l = 0; l < r->n_spec; ++l) {
      y[l] = z->sv[p + l * L->sz];
      E += (h[l] - tb->gas_const[l] * T) * y[l];
    }
    E -= oc_mixture_ve_energy(tb, T, y, NULL);
    E += z->sv[p + r->n_spec * L->sz];
    E *= rho;
  } else if (r->mixture) {
    double h[OC_MAX_SPEC];
    oc_enthalpy(tb, z->bv[p + 5 * L->sz], h);
    for (int l = 0; l < r->n_spec; ++l) E += h[l] * z->sv[p + l * L->sz];
    E *= rho;
    E -= z->bv[p + 4 * L->sz];
  } else {
    E *= rho;
    E += z->bv[p + 4 * L->sz] / (gamma_air - 1);
  }
  z->cv[p + 4 * L->sz] = E;
}

/* compute_cv_from_bv over the padded extent, src/FieldOperation.cuh:136-161 */
void oc_cv_from_bv(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  const lay_t L = layout(z);
  const int g = z->ngg;
#pragma omp parallel for schedule(static)
  for (int k = -g; k < z->mz + g; ++k)
    for (int j = -g; j < z->my + g; ++j)
      for (int i = -g; i < z->mx + g; ++i) cv_from_bv_point(z, &L, r, tb, IX(L, i, j, k));
}

/* update_physical_properties (1T), src/FieldOperation.cuh:183-247 */
void oc_update_properties(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  const lay_t L = layout(z);
  const int g = z->ngg, ns = r->n_spec;
#pragma omp parallel for schedule(static)
  for (int k = -g; k < z->mz + g; ++k)
    for (int j = -g; j < z->my + g; ++j)
      for (int i = -g; i < z->mx + g; ++i) {
        const long p = IX(L, i, j, k);
        const double T = z->bv[p + 5 * L.sz];
        const double V = norm3(z->bv[p + L.sz], z->bv[p + 2 * L.sz], z->bv[p + 3 * L.sz]);
        if (r->mixture) {
          double cp[OC_MAX_SPEC], y[OC_MAX_SPEC];
          oc_cp(tb, T, cp);
          double imw = 0, cpt = 0;
          for (int l = 0; l < ns; ++l) {
            y[l] = z->sv[p + l * L.sz];
            imw = fma(y[l], tb->imw[l], imw);
            cpt = fma(y[l], cp[l], cpt);
          }
          const double mw = 1 / dmax(imw, 1e-12);
          const double R = R_u * imw;
          double rd[OC_MAX_SPEC];
          if (r->two_temperature) {
            /* src/FieldOperation.cuh:207-229 */
            double cv_tr = 0, cp_tr = 0, r_mix = 0;
            oc_mixture_tr_energy(tb, T, y, &cv_tr, &cp_tr, &r_mix);
            const double gam = cp_tr / dmax(cv_tr, 1e-8);
            z->gamma[p] = gam;
            const double c = sqrt(dmax(gam * r_mix * T, 1e-12));
            z->acoustic[p] = c;
            z->mach[p] = V / c;
            if (z->tve[p] <= 0) z->tve[p] = T;
            oc_transport(tb, T, mw, y, cp, &z->mul[p], &z->kappa[p], rd);
            double cv_ve = 0;
            oc_mixture_ve_energy(tb, z->tve[p], y, &cv_ve);
            z->kappa_ve[p] = z->kappa[p] * cv_ve / dmax(cpt, 1e-8);
          } else {
            const double gam = cpt / (cpt - R);
            z->gamma[p] = gam;
            const double c = sqrt(gam * R * T);
            z->acoustic[p] = c;
            z->mach[p] = V / c;
            oc_transport(tb, T, mw, y, cp, &z->mul[p], &z->kappa[p], rd);
          }
          for (int l = 0; l < ns; ++l) z->rhoD[p + l * L.sz] = rd[l];
        } else {
          const double c_temp = gamma_air * R_u / mw_air;
          z->mul[p] = sutherland(T);
          z->mach[p] = V / sqrt(c_temp * T);
        }
      }
}

void oc_zero_dq(oc_zone *z, const oc_run *r) { memset(z->dq, 0, sizeof(double) * DQSZ(z) * r->n_var); }

/* compute_temperature_and_pressure, src/FieldOperation.cu:4-94 */
static void temperature_and_pressure(oc_zone *z, const lay_t *L, const oc_run *r, const oc_tables *tb, long p,
                                     double total_energy) {
  const int ns = tb->n_spec;
  const double density = z->bv[p];
  double gas_const = 0;
  for (int l = 0; l < ns; ++l) gas_const += z->sv[p + l * L->sz] * tb->gas_const[l];
  const double u = z->bv[p + L->sz], v = z->bv[p + 2 * L->sz], w = z->bv[p + 3 * L->sz];
  const double e = total_energy / density - 0.5 * (u * u + v * v + w * w);
  if (r->two_temperature) {
    /* Tve from E_ve, then T from e - E_ve = E_tr(T) (src/FieldOperation.cu:24-64) */
    double y[OC_MAX_SPEC];
    for (int l = 0; l < ns; ++l) y[l] = z->sv[p + l * L->sz];
    double eve = dmax(z->sv[p + ns * L->sz], 0.0);
    double tve = z->tve[p];
    if (tve <= 0) tve = dmax(z->bv[p + 5 * L->sz], 1.0);
    tve = oc_invert_tve(tb, eve, y, tve);
    const double max_tve = dmax(1.0e7, 1000.0 * dmax(z->bv[p + 5 * L->sz], 1.0));
    if (!isfinite(eve) || !isfinite(tve) || tve <= 0.0 || tve > max_tve) {
      tve = dmax(z->bv[p + 5 * L->sz], 1.0);
      eve = oc_mixture_ve_energy(tb, tve, y, NULL);
      z->sv[p + ns * L->sz] = eve;
      z->cv[p + (5 + ns) * L->sz] = density * eve;
    }
    z->tve[p] = tve;
    double err = 1, t = dmax(z->bv[p + 5 * L->sz], 1.0);
    int iter = 0;
    while (err > 1e-5 && iter++ < 1000) {
      double cv_tr = 0;
      const double e_tr = oc_mixture_tr_energy(tb, t, y, &cv_tr, NULL, NULL);
      const double t1 = dmax(1.0, t - (e_tr - (e - eve)) / dmax(cv_tr, 1e-8));
      err = fabs(1 - t1 / t);
      t = t1;
    }
    z->bv[p + 5 * L->sz] = t;
    z->bv[p + 4 * L->sz] = density * t * gas_const;
    return;
  }
  double err = 1, t = z->bv[p + 5 * L->sz];
  int iter = 0;
  double h[OC_MAX_SPEC], cp[OC_MAX_SPEC];
  while (err > 1e-5 && iter++ < 1000) {
    oc_enthalpy_and_cp(tb, t, h, cp);
    double cp_tot = 0, hh = 0;
    for (int l = 0; l < ns; ++l) {
      cp_tot += cp[l] * z->sv[p + l * L->sz];
      hh += h[l] * z->sv[p + l * L->sz];
    }
    const double e_t = hh - gas_const * t;
    const double cvv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cvv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  z->bv[p + 5 * L->sz] = t;
  z->bv[p + 4 * L->sz] = z->bv[p] * t * gas_const;
}

/* The primitive recovery after an update of cv at point p: velocities, the
 * species clipped at 0 and rescaled to sum to rho, then T and p
 * (src/RK.cuh:300-366; the same block ends update_cv_and_bv_rk_wu and
 * wu_zeroDReaction, src/StrangSplitting.cu:533-565, 193-227). */
static void recover_point(oc_zone *z, const lay_t *Lp, const oc_run *r, const oc_tables *tb, long p);

/* update_cv_and_bv_rk (1T), src/RK.cuh:277-367; with cn (Wu splitting)
 * the increment is dq - Cn (update_cv_and_bv_rk_wu, src/StrangSplitting.cu:508-566) */
static void update_rk_impl(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk, const double *cn) {
  const lay_t L = layout(z);
  const int nv = r->n_var;
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k), q = DQX(z, i, j, k);
        double *cv = z->cv;
        const double dt_div_jac = dt / z->jac[p];
        if (cn) {
          for (int l = 0; l < nv; ++l)
            cv[p + l * L.sz] = rk_a[rk] * z->qn[p + l * L.sz] + rk_b[rk] * cv[p + l * L.sz] +
                               rk_c[rk] * dt_div_jac * (z->dq[q + l * dsz] - cn[q + l * dsz]);
        } else {
          for (int l = 0; l < nv; ++l)
            cv[p + l * L.sz] = rk_a[rk] * z->qn[p + l * L.sz] + rk_b[rk] * cv[p + l * L.sz] +
                               rk_c[rk] * dt_div_jac * z->dq[q + l * dsz];
        }
        if (z->mz == 1) cv[p + 3 * L.sz] = 0;
        recover_point(z, &L, r, tb, p);
      }
}

void oc_update_rk(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk) {
  update_rk_impl(z, r, tb, dt, rk, NULL);
}

void oc_update_rk_wu(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk) {
  update_rk_impl(z, r, tb, dt, rk, z->cn);
}

static void recover_point(oc_zone *z, const lay_t *Lp, const oc_run *r, const oc_tables *tb, long p) {
  const lay_t L = *Lp;
  const int ns = r->n_spec;
  double *cv = z->cv, *bv = z->bv, *sv = z->sv;
  bv[p] = cv[p];
  const double density_inv = 1.0 / cv[p];
  bv[p + L.sz] = cv[p + L.sz] * density_inv;
  bv[p + 2 * L.sz] = cv[p + 2 * L.sz] * density_inv;
  bv[p + 3 * L.sz] = cv[p + 3 * L.sz] * density_inv;
  const double V2 =
      bv[p + L.sz] * bv[p + L.sz] + bv[p + 2 * L.sz] * bv[p + 2 * L.sz] + bv[p + 3 * L.sz] * bv[p + 3 * L.sz];
  if (r->mixture) {
    double sum = 0;
    for (int l = 0; l < ns; ++l) {
      if (cv[p + (5 + l) * L.sz] < 0)
        cv[p + (5 + l) * L.sz] = 0;
      else
        sum += cv[p + (5 + l) * L.sz];
    }
    if (r->two_temperature) cv[p + (5 + ns) * L.sz] = dmax(cv[p + (5 + ns) * L.sz], 0.0);
    sum = 1.0 / sum;
    const double ratio = cv[p] * sum;
    for (int l = 0; l < ns; ++l) {
      sv[p + l * L.sz] = cv[p + (5 + l) * L.sz] * sum;
      cv[p + (5 + l) * L.sz] *= ratio;
    }
  }
  for (int l = ns; l < r->n_scalar; ++l) sv[p + l * L.sz] = cv[p + (5 + l) * L.sz] * density_inv;
  if (r->mixture) {
    temperature_and_pressure(z, &L, r, tb, p, cv[p + 4 * L.sz]);
  } else {
    bv[p + 4 * L.sz] = (gamma_air - 1) * (cv[p + 4 * L.sz] - 0.5 * bv[p] * V2);
    bv[p + 5 * L.sz] = bv[p + 4 * L.sz] * mw_air * density_inv / R_u;
  }
}

/* apply_vt_relaxation_1_point over the interior, src/RK.cuh:20-28,
 * src/FieldOperation.cuh:60-88: the exact exponential Landau-Teller relaxation
 * of the vibrational energy over the stage's dt */
void oc_vt_relaxation(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt_stage) {
  if (!r->two_temperature || !r->vt_relaxation || dt_stage <= 0.0) return;
  const lay_t L = layout(z);
  const int ns = r->n_spec;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k);
        double *cv = z->cv;
        const double density = cv[p];
        if (density <= 0.0) continue;
        const double tve = z->tve[p] > 0.0 ? z->tve[p] : dmax(z->bv[p + 5 * L.sz], 1.0);
        double y[OC_MAX_SPEC];
        for (int l = 0; l < ns; ++l) y[l] = z->sv[p + l * L.sz];
        double ev_eq = 0, tau = 0;
        oc_vt_source(tb, density, z->bv[p + 5 * L.sz], tve, y, &ev_eq, &tau);
        if (tau >= 1e29 || !isfinite(tau)) continue;
        const double rho_eve_old = dmax(cv[p + (5 + ns) * L.sz], 0.0);
        const double alpha = exp(-dt_stage / dmax(tau, 1e-30));
        const double ev_old = oc_mixture_vib_energy(tb, tve, y, NULL);
        const double ev_new = ev_old * alpha + ev_eq * (1.0 - alpha);
        const double rho_eve_new = rho_eve_old + density * (ev_new - ev_old);
        cv[p + (5 + ns) * L.sz] = dmax(rho_eve_new, 0.0);
        z->sv[p + ns * L.sz] = cv[p + (5 + ns) * L.sz] / density;
        z->tve[p] = oc_invert_tve(tb, z->sv[p + ns * L.sz], y, tve);
        temperature_and_pressure(z, &L, r, tb, p, cv[p + 4 * L.sz]);
      }
}

/* finite_rate_chemistry (explicit), src/FiniteRateChem.cu:1406-1488 */
void oc_chemistry(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  const lay_t L = layout(z);
  const int ns = r->n_spec;
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k), q = DQX(z, i, j, k);
        double y[OC_MAX_SPEC], omega[OC_MAX_SPEC], omega_d[OC_MAX_SPEC];
        const double density = z->bv[p];
        for (int l = 0; l < ns; ++l) y[l] = z->sv[p + l * L.sz];
        const double t = z->bv[p + 5 * L.sz];
        const double tve = r->two_temperature ? dmax(z->tve[p], 1.0) : t;
        oc_chem_source(tb, density, t, tve, y, omega, omega_d);
        double time_scale = 1e+6;
        for (int l = 0; l < ns; ++l) {
          z->dq[q + (l + 5) * dsz] += z->jac[p] * omega[l];
          if (y[l] > 1e-25 && fabs(omega_d[l]) > 1e-25)
            time_scale = dmin(time_scale, fabs(density * y[l] / omega_d[l]));
        }
        if (r->two_temperature) {
          /* compute_two_temperature_chemical_source, src/FiniteRateChem.cu:19-31 */
          double tv = z->tve[p];
          if (tv <= 0.0) tv = dmax(t, 1.0);
          double src = 0.0;
          for (int l = 0; l < ns; ++l) src += omega[l] * oc_ve_energy(tb, l, tv);
          z->dq[q + (5 + ns) * dsz] += z->jac[p] * src;
        }
        z->rts[p] = time_scale;
      }
}

/* ------------------------------------------------------------------------ */
/* WENO5-JS / WENO7-Z, src/WENO.cu:3033-3196                                 */
/* ------------------------------------------------------------------------ */
double oc_weno5(const double *vp, const double *vm, double eps, int if_shock) {
  const double one6th = 1.0 / 6;
  if (if_shock) {
    double v0 = one6th * (2 * vp[2] + 5 * vp[3] - vp[4]);
    double v1 = one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3]);
    double v2 = one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2]);
    const double thirteen12th = 13.0 / 12;
    double beta0 = thirteen12th * (vp[2] + vp[4] - 2 * vp[3]) * (vp[2] + vp[4] - 2 * vp[3]) +
                   0.25 * (3 * vp[2] - 4 * vp[3] + vp[4]) * (3 * vp[2] - 4 * vp[3] + vp[4]);
    double beta1 = thirteen12th * (vp[1] + vp[3] - 2 * vp[2]) * (vp[1] + vp[3] - 2 * vp[2]) +
                   0.25 * (vp[1] - vp[3]) * (vp[1] - vp[3]);
    double beta2 = thirteen12th * (vp[0] + vp[2] - 2 * vp[1]) * (vp[0] + vp[2] - 2 * vp[1]) +
                   0.25 * (vp[0] - 4 * vp[1] + 3 * vp[2]) * (vp[0] - 4 * vp[1] + 3 * vp[2]);
    double a0 = 0.3 / ((eps + beta0) * (eps + beta0));
    double a1 = 0.6 / ((eps + beta1) * (eps + beta1));
    double a2 = 0.1 / ((eps + beta2) * (eps + beta2));
    const double fPlus = (a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2);

    v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
    v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
    v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
    beta0 = thirteen12th * (vm[2] + vm[4] - 2 * vm[3]) * (vm[2] + vm[4] - 2 * vm[3]) +
            0.25 * (3 * vm[2] - 4 * vm[3] + vm[4]) * (3 * vm[2] - 4 * vm[3] + vm[4]);
    beta1 = thirteen12th * (vm[1] + vm[3] - 2 * vm[2]) * (vm[1] + vm[3] - 2 * vm[2]) +
            0.25 * (vm[1] - vm[3]) * (vm[1] - vm[3]);
    beta2 = thirteen12th * (vm[0] + vm[2] - 2 * vm[1]) * (vm[0] + vm[2] - 2 * vm[1]) +
            0.25 * (vm[0] - 4 * vm[1] + 3 * vm[2]) * (vm[0] - 4 * vm[1] + 3 * vm[2]);
    a0 = 0.1 / ((eps + beta0) * (eps + beta0));
    a1 = 0.6 / ((eps + beta1) * (eps + beta1));
    a2 = 0.3 / ((eps + beta2) * (eps + beta2));
    const double fMinus = (a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2);
    return fPlus + fMinus;
  }
  double v0 = one6th * (2 * vp[2] + 5 * vp[3] - vp[4]);
  double v1 = one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3]);
  double v2 = one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2]);
  const double fPlus = 0.3 * v0 + 0.6 * v1 + 0.1 * v2;
  v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
  v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
  v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
  const double fMinus = 0.1 * v0 + 0.6 * v1 + 0.3 * v2;
  return fPlus + fMinus;
}

static double weno7_beta(double s1, double s2, double s3) {
  const double d12 = 13.0 / 12.0, d13 = 1043.0 / 960, d14 = 1.0 / 12;
  return s1 * s1 + d12 * s2 * s2 + d13 * s3 * s3 + d14 * s1 * s3;
}

double oc_weno7(const double *vp, const double *vm, double eps, int if_shock) {
  const double one6th = 1.0 / 6, one12th = 1.0 / 12;
  const double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35;
  if (if_shock) {
    double s1 = one6th * (-2 * vp[0] + 9 * vp[1] - 18 * vp[2] + 11 * vp[3]);
    double s2 = -vp[0] + 4 * vp[1] - 5 * vp[2] + 2 * vp[3];
    double s3 = -vp[0] + 3 * vp[1] - 3 * vp[2] + vp[3];
    double beta0 = weno7_beta(s1, s2, s3);
    s1 = one6th * (vp[1] - 6 * vp[2] + 3 * vp[3] + 2 * vp[4]);
    s2 = vp[2] - 2 * vp[3] + vp[4];
    s3 = -vp[1] + 3 * vp[2] - 3 * vp[3] + vp[4];
    double beta1 = weno7_beta(s1, s2, s3);
    s1 = one6th * (-2 * vp[2] - 3 * vp[3] + 6 * vp[4] - vp[5]);
    s3 = -vp[2] + 3 * vp[3] - 3 * vp[4] + vp[5];
    double beta2 = weno7_beta(s1, s2, s3);
    s1 = one6th * (-11 * vp[3] + 18 * vp[4] - 9 * vp[5] + 2 * vp[6]);
    s2 = 2 * vp[3] - 5 * vp[4] + 4 * vp[5] - vp[6];
    s3 = -vp[3] + 3 * vp[4] - 3 * vp[5] + vp[6];
    double beta3 = weno7_beta(s1, s2, s3);
    double tau = beta0 + 3 * beta1 - 3 * beta2 - beta3;
    tau *= tau;
    double a0 = c0 + c0 * tau / ((eps + beta0) * (eps + beta0));
    double a1 = c1 + c1 * tau / ((eps + beta1) * (eps + beta1));
    double a2 = c2 + c2 * tau / ((eps + beta2) * (eps + beta2));
    double a3 = c3 + c3 * tau / ((eps + beta3) * (eps + beta3));
    double v0 = -3 * vp[0] + 13 * vp[1] - 23 * vp[2] + 25 * vp[3];
    double v1 = vp[1] - 5 * vp[2] + 13 * vp[3] + 3 * vp[4];
    double v2 = -vp[2] + 7 * vp[3] + 7 * vp[4] - vp[5];
    double v3 = 3 * vp[3] + 13 * vp[4] - 5 * vp[5] + vp[6];
    const double fPlus = one12th * (a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3);

    s1 = one6th * (-2 * vm[6] + 9 * vm[5] - 18 * vm[4] + 11 * vm[3]);
    s2 = -vm[6] + 4 * vm[5] - 5 * vm[4] + 2 * vm[3];
    s3 = -vm[6] + 3 * vm[5] - 3 * vm[4] + vm[3];
    beta0 = weno7_beta(s1, s2, s3);
    s1 = one6th * (vm[5] - 6 * vm[4] + 3 * vm[3] + 2 * vm[2]);
    s2 = vm[4] - 2 * vm[3] + vm[2];
    s3 = -vm[5] + 3 * vm[4] - 3 * vm[3] + vm[2];
    beta1 = weno7_beta(s1, s2, s3);
    s1 = one6th * (-2 * vm[4] - 3 * vm[3] + 6 * vm[2] - vm[1]);
    s3 = -vm[4] + 3 * vm[3] - 3 * vm[2] + vm[1];
    beta2 = weno7_beta(s1, s2, s3);
    s1 = one6th * (-11 * vm[3] + 18 * vm[2] - 9 * vm[1] + 2 * vm[0]);
    s2 = 2 * vm[3] - 5 * vm[2] + 4 * vm[1] - vm[0];
    s3 = -vm[3] + 3 * vm[2] - 3 * vm[1] + vm[0];
    beta3 = weno7_beta(s1, s2, s3);
    tau = beta0 + 3 * beta1 - 3 * beta2 - beta3;
    tau *= tau;
    a0 = c0 + c0 * tau / ((eps + beta0) * (eps + beta0));
    a1 = c1 + c1 * tau / ((eps + beta1) * (eps + beta1));
    a2 = c2 + c2 * tau / ((eps + beta2) * (eps + beta2));
    a3 = c3 + c3 * tau / ((eps + beta3) * (eps + beta3));
    v0 = -3 * vm[6] + 13 * vm[5] - 23 * vm[4] + 25 * vm[3];
    v1 = vm[5] - 5 * vm[4] + 13 * vm[3] + 3 * vm[2];
    v2 = -vm[4] + 7 * vm[3] + 7 * vm[2] - vm[1];
    v3 = 3 * vm[3] + 13 * vm[2] - 5 * vm[1] + vm[0];
    const double fMinus = one12th * (a0 * v0 + a1 * v1 + a2 * v2 + a3 * v3) / (a0 + a1 + a2 + a3);
    return fPlus + fMinus;
  }
  double v3 = 3 * vp[3] + 13 * vp[4] - 5 * vp[5] + vp[6];
  double v2 = -vp[2] + 7 * vp[3] + 7 * vp[4] - vp[5];
  double v1 = vp[1] - 5 * vp[2] + 13 * vp[3] + 3 * vp[4];
  double v0 = -3 * vp[0] + 13 * vp[1] - 23 * vp[2] + 25 * vp[3];
  const double fPlus = one12th * (c0 * v0 + c1 * v1 + c2 * v2 + c3 * v3);
  v0 = -3 * vm[6] + 13 * vm[5] - 23 * vm[4] + 25 * vm[3];
  v1 = vm[5] - 5 * vm[4] + 13 * vm[3] + 3 * vm[2];
  v2 = -vm[4] + 7 * vm[3] + 7 * vm[2] - vm[1];
  v3 = 3 * vm[3] + 13 * vm[2] - 5 * vm[1] + vm[0];
  const double fMinus = one12th * (c0 * v0 + c1 * v1 + c2 * v2 + c3 * v3);
  return fPlus + fMinus;
}

/* ------------------------------------------------------------------------ */
/* characteristic WENO face flux                                              */
/* compute_convective_term_weno_{x,y,z}, char branch, src/WENO.cu:1090-1457, */
/* 1818-2193, 2483-2857; Air helpers weno_ch_air_x/yz src/WENO.cu:10-244,    */
/* 414-654; mixture thermo src/Thermo.cuh:49-110                             */
/* ------------------------------------------------------------------------ */
#define OC_MAXV (5 + OC_MAX_SPEC)

/* stencil point data: Q[nv], F0..F6 */
typedef struct {
  double Q[OC_MAXV];
  double F[7];
} spt_t;

static void load_stencil_point(const oc_zone *z, const lay_t *L, const oc_run *r, long p, int dir, spt_t *s) {
  const long sz = L->sz;
  const double q0 = z->cv[p], q1 = z->cv[p + sz], q2 = z->cv[p + 2 * sz], q3 = z->cv[p + 3 * sz];
  for (int l = 0; l < r->n_var; ++l) s->Q[l] = z->cv[p + l * sz];
  const double m0 = z->metric[p + (3 * dir + 0) * sz];
  const double m1 = z->metric[p + (3 * dir + 1) * sz];
  const double m2 = z->metric[p + (3 * dir + 2) * sz];
  const double Uk = (q1 * m0 + q2 * m1 + q3 * m2) / q0;
  const double pk = z->bv[p + 4 * sz];
  s->F[0] = Uk;
  s->F[1] = fma(q1, Uk, pk * m0);
  s->F[2] = fma(q2, Uk, pk * m1);
  s->F[3] = fma(q3, Uk, pk * m2);
  s->F[4] = pk;
  double cGradK = norm3(m0, m1, m2);
  if (r->mixture)
    cGradK *= z->acoustic[p];
  else
    cGradK *= sqrt(gamma_air * R_AIR * z->bv[p + 5 * sz]);
  s->F[5] = cGradK;
  s->F[6] = z->jac[p];
}

/* compute_mixture_characteristic_thermo, src/Thermo.cuh:49-110: the
 * translational-rotational part of the equilibrium thermodynamics (E_ve(T)
 * and its c_v removed; nothing to remove with one temperature) and, with two
 * temperatures, the E_ve row of the scalar coefficients */
static void mixture_char_thermo(const oc_tables *tb, const oc_run *r, double t, const double *y,
                                double *alpha_out, double *ecoef_out, double *gamma_out, double *c_out) {
  const double t_eval = dmax(t, 1.0);
  double h_eq[OC_MAX_SPEC], cp_eq[OC_MAX_SPEC], h_tr[OC_MAX_SPEC];
  oc_enthalpy_and_cp(tb, t_eval, h_eq, cp_eq);
  double r_local = 0, cp_local = 0, cv_local = 0;
  for (int l = 0; l < tb->n_spec; ++l) {
    const double yi = y[l];
    const double eve_eq = oc_ve_energy(tb, l, t_eval), cv_ve_eq = oc_ve_cv(tb, l, t_eval);
    h_tr[l] = h_eq[l] - eve_eq;
    const double cp_tr = cp_eq[l] - cv_ve_eq;
    r_local += yi * tb->gas_const[l];
    cp_local += yi * cp_tr;
    cv_local += yi * (cp_tr - tb->gas_const[l]);
  }
  const double gamma_local = cp_local / dmax(cv_local, 1e-8);
  const double c_local = sqrt(dmax(gamma_local * r_local * t_eval, 1e-12));
  *gamma_out = gamma_local;
  *c_out = c_local;
  for (int l = 0; l < r->n_scalar_transported; ++l) {
    alpha_out[l] = 0.0;
    ecoef_out[l] = 0.0;
  }
  const double gm1 = gamma_local - 1.0;
  const double c2 = dmax(c_local * c_local, 1e-30);
  const double inv_gm1 = 1.0 / dmax(gm1, 1e-12);
  for (int l = 0; l < tb->n_spec; ++l) {
    const double alpha = (gamma_local * tb->gas_const[l] * t_eval - gm1 * h_tr[l]) / c2;
    alpha_out[l] = alpha;
    ecoef_out[l] = -alpha * c2 * inv_gm1;
  }
  if (r->two_temperature) {
    alpha_out[tb->n_spec] = -gm1 / c2;
    ecoef_out[tb->n_spec] = 1.0;
  }
}

/* The Roe-averaged state of a face and its characteristic basis
 * (src/WENO.cu:1146-1219 x, 1872-1946 y, 2537-2611 z): Roe weights,
 * mixture thermodynamics at the Roe temperature, the face normal k and the
 * tangents n, q (branch on |k_x| > 0.8). */
typedef struct {
  int mixture, nsc;
  double rlc, rrc, um, vm, wm, cm, gm1, temp2, temp3;
  double kx, ky, kz, nx, ny, nz, qx, qy, qz, Uk_bar, Un, Uq, Hroe;
  double svm[OC_MAX_SPEC + 4], alpha[OC_MAX_SPEC + 4], ecoef[OC_MAX_SPEC + 4];
} char_state_t;

static void char_state(const oc_run *r, const oc_tables *tb, const spt_t *Lp, const spt_t *Rp, double kx, double ky,
                       double kz, char_state_t *st) {
  const int nv = r->n_var;
  memset(st, 0, sizeof(*st));
  st->mixture = r->mixture;
  st->nsc = r->mixture ? nv - 5 : 0;
  const double rho_l = Lp->Q[0], rho_r = Rp->Q[0];
  double temp1 = sqrt(rho_l * rho_r);
  const double rlc = 1 / (rho_l + temp1);
  const double rrc = 1 / (temp1 + rho_r);
  st->rlc = rlc;
  st->rrc = rrc;
  const double um = rlc * Lp->Q[1] + rrc * Rp->Q[1];
  const double vm = rlc * Lp->Q[2] + rrc * Rp->Q[2];
  const double wm = rlc * Lp->Q[3] + rrc * Rp->Q[3];
  double gamma, cm, gm1;
  if (r->mixture) {
    for (int l = 0; l < st->nsc; ++l) st->svm[l] = rlc * Lp->Q[l + 5] + rrc * Rp->Q[l + 5];
    double r_mix_guess = 0.0;
    for (int l = 0; l < r->n_spec; ++l) r_mix_guess += st->svm[l] * tb->gas_const[l];
    const double Troe = (rlc * Lp->F[4] + rrc * Rp->F[4]) / dmax(r_mix_guess, 1e-12);
    mixture_char_thermo(tb, r, Troe, st->svm, st->alpha, st->ecoef, &gamma, &cm);
    gm1 = gamma - 1;
  } else {
    const double Troe = (rlc * Lp->F[4] + rrc * Rp->F[4]) / R_AIR;
    cm = sqrt(gamma_air * R_AIR * Troe);
    gm1 = gamma_air - 1;
  }
  temp1 = 1.0 / norm3(kx, ky, kz);
  kx *= temp1;
  ky *= temp1;
  kz *= temp1;
  double nx = 0, ny = 0, nz = 0, qx, qy, qz;
  if (fabs(kx) > 0.8) {
    nx = -kz;
    nz = kx;
  } else {
    ny = kz;
    nz = -ky;
  }
  temp1 = 1.0 / norm3(nx, ny, nz);
  nx *= temp1;
  ny *= temp1;
  nz *= temp1;
  qx = ky * nz - kz * ny;
  qy = kz * nx - kx * nz;
  qz = kx * ny - ky * nx;
  temp1 = 1.0 / norm3(qx, qy, qz);
  qx *= temp1;
  qy *= temp1;
  qz *= temp1;
  st->um = um;
  st->vm = vm;
  st->wm = wm;
  st->cm = cm;
  st->gm1 = gm1;
  st->kx = kx;
  st->ky = ky;
  st->kz = kz;
  st->nx = nx;
  st->ny = ny;
  st->nz = nz;
  st->qx = qx;
  st->qy = qy;
  st->qz = qz;
  st->Uk_bar = kx * um + ky * vm + kz * wm;
  st->Un = nx * um + ny * vm + nz * wm;
  st->Uq = qx * um + qy * vm + qz * wm;
  st->temp2 = 1.0 / (cm * cm);
  st->temp3 = 0.5 * gm1 * (um * um + vm * vm + wm * wm);
  st->Hroe = rlc * (Lp->Q[4] + Lp->F[4]) + rrc * (Rp->Q[4] + Rp->F[4]);
}

/* Left eigenvector row l < 5 on (rho, rho u, rho v, rho w, rho E) and the
 * factor tf of its scalar part tf * sum_n alpha_n (rho Y_n),
 * src/WENO.cu:1223-1330 (docs/two_temperature_weno_characteristic_derivation.md). */
static double char_row(const char_state_t *st, int l, double *Lr) {
  const double temp2 = st->temp2, temp3 = st->temp3, gm1 = st->gm1, cm = st->cm;
  const double um = st->um, vm = st->vm, wm = st->wm;
  switch (l) {
    case 0:
      Lr[0] = (temp3 + st->Uk_bar * cm) * temp2 * 0.5;
      Lr[1] = -(gm1 * um + st->kx * cm) * temp2 * 0.5;
      Lr[2] = -(gm1 * vm + st->ky * cm) * temp2 * 0.5;
      Lr[3] = -(gm1 * wm + st->kz * cm) * temp2 * 0.5;
      Lr[4] = gm1 * temp2 * 0.5;
      return 0.5;
    case 1:
      Lr[0] = -st->Un;
      Lr[1] = st->nx;
      Lr[2] = st->ny;
      Lr[3] = st->nz;
      Lr[4] = 0;
      return 0;
    case 2:
      Lr[0] = -st->Uq;
      Lr[1] = st->qx;
      Lr[2] = st->qy;
      Lr[3] = st->qz;
      Lr[4] = 0;
      return 0;
    case 3:
      Lr[0] = 1 - temp3 * temp2;
      Lr[1] = gm1 * um * temp2;
      Lr[2] = gm1 * vm * temp2;
      Lr[3] = gm1 * wm * temp2;
      Lr[4] = -gm1 * temp2;
      return -1;
    default:
      Lr[0] = (temp3 - st->Uk_bar * cm) * temp2 * 0.5;
      Lr[1] = -(gm1 * um - st->kx * cm) * temp2 * 0.5;
      Lr[2] = -(gm1 * vm - st->ky * cm) * temp2 * 0.5;
      Lr[3] = -(gm1 * wm - st->kz * cm) * temp2 * 0.5;
      Lr[4] = gm1 * temp2 * 0.5;
      return 0.5;
  }
}

/* sum_n alpha_n Q[5 + n], the scalar part shared by the acoustic and entropy rows */
static double char_beta(const char_state_t *st, const double *Q) {
  double sp = 0;
  for (int n = 0; n < st->nsc; ++n) sp = fma(st->alpha[n], Q[5 + n], sp);
  return sp;
}

/* Component l of L Q (the characteristic projection of a conservative vector) */
static double char_project(const char_state_t *st, int l, const double *Lr, double tf, const double *Q,
                           double sum_beta) {
  if (l >= 5) return Q[l] - st->svm[l - 5] * Q[0];
  double LQm = Lr[0] * Q[0] + Lr[1] * Q[1] + Lr[2] * Q[2] + Lr[3] * Q[3] + Lr[4] * Q[4];
  if (st->mixture) LQm += tf * sum_beta;
  return LQm;
}

/* back-projection R fChar, src/WENO.cu:1439-1457 */
static void char_back(const char_state_t *st, const double *fChar, double *fc) {
  const double cm = st->cm;
  const double temp1 = fChar[0] + fChar[3] + fChar[4];
  const double t3b = fChar[0] - fChar[4];
  fc[0] = temp1;
  fc[1] = st->um * temp1 - st->kx * cm * t3b + st->nx * fChar[1] + st->qx * fChar[2];
  fc[2] = st->vm * temp1 - st->ky * cm * t3b + st->ny * fChar[1] + st->qy * fChar[2];
  fc[3] = st->wm * temp1 - st->kz * cm * t3b + st->nz * fChar[1] + st->qz * fChar[2];
  fc[4] = st->Hroe * temp1 - st->Uk_bar * cm * t3b + st->Un * fChar[1] + st->Uq * fChar[2] -
          cm * cm / st->gm1 * fChar[3];
  if (st->mixture) {
    double t2 = 0;
    for (int l = 0; l < st->nsc; ++l) {
      fc[5 + l] = st->svm[l] * temp1 + fChar[l + 5];
      t2 += st->ecoef[l] * fChar[l + 5];
    }
    fc[4] += t2;
  }
}

/* One face flux between stencil points (window of npt = 6 or 8 points,
 * the face lies between window index ic and ic+1). */
static void char_weno_face(const oc_run *r, const oc_tables *tb, const spt_t *S, int ic, double kx, double ky,
                           double kz, double eps_ref, int if_shock, double *fc) {
  const int w7 = r->inviscid_scheme == 72;
  const int npt = w7 ? 8 : 6;
  const int base = w7 ? ic - 3 : ic - 2;
  char_state_t st;
  char_state(r, tb, &S[ic], &S[ic + 1], kx, ky, kz, &st);

  double max_lambda = 0;
  for (int m = 0; m < npt; ++m) {
    const double lam = fabs(S[base + m].F[0]) + S[base + m].F[5];
    if (lam > max_lambda) max_lambda = lam;
  }
  double sumBetaQ[8] = {0}, sumBetaF[8] = {0};
  if (r->mixture) {
    for (int m = 0; m < npt; ++m) {
      const spt_t *P = &S[base + m];
      double sm = 0;
      for (int n = 0; n < st.nsc; ++n) sm = fma(st.alpha[n], P->Q[5 + n] * P->F[0], sm);
      sumBetaQ[m] = char_beta(&st, P->Q);
      sumBetaF[m] = sm;
    }
  }

  double fChar[OC_MAXV];
  for (int l = 0; l < 5; ++l) {
    double Lr[5];
    const double tf = char_row(&st, l, Lr);
    double vPlus[7] = {0}, vMinus[7] = {0};
    for (int m = 0; m < npt; ++m) {
      const spt_t *P = &S[base + m];
      double LFm = Lr[0] * (P->Q[0] * P->F[0]) + Lr[1] * P->F[1] + Lr[2] * P->F[2] + Lr[3] * P->F[3] +
                   Lr[4] * ((P->Q[4] + P->F[4]) * P->F[0]);
      if (r->mixture) LFm += tf * sumBetaF[m];
      const double LQm = char_project(&st, l, Lr, tf, P->Q, sumBetaQ[m]);
      if (m < npt - 1) vPlus[m] = 0.5 * P->F[6] * (LFm + max_lambda * LQm);
      if (m > 0) vMinus[m - 1] = 0.5 * P->F[6] * (LFm - max_lambda * LQm);
    }
    fChar[l] = w7 ? oc_weno7(vPlus, vMinus, eps_ref, if_shock) : oc_weno5(vPlus, vMinus, eps_ref, if_shock);
  }
  for (int l = 0; l < st.nsc; ++l) {
    double vPlus[7], vMinus[7];
    for (int m = 0; m < npt; ++m) {
      const spt_t *P = &S[base + m];
      const double LQm = char_project(&st, 5 + l, NULL, 0, P->Q, 0);
      const double LFm = P->F[0] * LQm;
      if (m < npt - 1) vPlus[m] = 0.5 * P->F[6] * (LFm + max_lambda * LQm);
      if (m > 0) vMinus[m - 1] = 0.5 * P->F[6] * (LFm - max_lambda * LQm);
    }
    fChar[5 + l] = w7 ? oc_weno7(vPlus, vMinus, eps_ref, if_shock) : oc_weno5(vPlus, vMinus, eps_ref, if_shock);
  }
  char_back(&st, fChar, fc);
}

/* The characteristic basis of char_weno_face for the face between two
 * states, as matrices (row-major, n_var x n_var): left_m = L (the projection
 * char_project applied to unit vectors), right_m = R (char_back applied to
 * unit vectors).  q_l/q_r are conservative vectors, p_l/p_r pressures. */
void oc_char_eigen(const oc_run *r, const oc_tables *tb, const double *q_l, double p_l, const double *q_r, double p_r,
                   double kx, double ky, double kz, double *left_m, double *right_m) {
  const int nv = r->n_var;
  spt_t Lp, Rp;
  memset(&Lp, 0, sizeof(Lp));
  memset(&Rp, 0, sizeof(Rp));
  for (int l = 0; l < nv; ++l) {
    Lp.Q[l] = q_l[l];
    Rp.Q[l] = q_r[l];
  }
  Lp.F[4] = p_l;
  Rp.F[4] = p_r;
  char_state_t st;
  char_state(r, tb, &Lp, &Rp, kx, ky, kz, &st);
  for (int c = 0; c < nv; ++c) {
    double e[OC_MAXV] = {0}, fc[OC_MAXV] = {0};
    e[c] = 1.0;
    const double sb = char_beta(&st, e);
    for (int l = 0; l < nv; ++l) {
      double Lr[5] = {0};
      const double tf = l < 5 ? char_row(&st, l, Lr) : 0;
      left_m[l * nv + c] = char_project(&st, l, Lr, tf, e, sb);
    }
    char_back(&st, e, fc);
    for (int l = 0; l < nv; ++l) right_m[l * nv + c] = fc[l];
  }
}

/* if_shock of a face, src/WENO.cu:839-849 (1568-1578, 2233-2243): the
 * sensor anywhere in [left - ngg + 1, left + ngg] above the threshold, or no
 * hybrid threshold at all.  Sensor ghosts are never written and stay 0. */
static int face_if_shock(const oc_zone *z, const oc_run *r, long p, long stride) {
  if (!(r->sensor_threshold > 1e-10)) return 1;
  for (int ii = -z->ngg + 1; ii <= z->ngg; ++ii)
    if (z->shock_sensor[p + ii * stride] > r->sensor_threshold) return 1;
  return 0;
}

/* The time step the positivity limiter scales the fluxes with: DParameter::dt
 * when the step is fixed, else dt_local of the face's left node read with the
 * interior array's flat index (src/WENO.cu:992-996).  Face -1 of the first row
 * has a negative flat index (a read before the array in the reference); it is
 * clamped to 0 here. */
static double pp_time_step(const oc_zone *z, const oc_run *r, int i, int j, int k) {
  if (r->pp_dt > 0) return r->pp_dt;
  long q = DQX(z, i, j, k);
  if (q < 0) q = 0;
  return z->dt_local[q];
}

/* One species (or E_ve) flux through the positivity limiter,
 * src/WENO.cu:1017-1083: blend with the first-order flux f1 where the flux
 * would drive rhoY_left - dt F or rhoY_right + dt F negative.  f1 enters as
 * the two halves (left part, right part) the caller forms. */
static double pp_limit(double fc, double alphaL, double alphaR, double rhoYL, double rhoYR, double dt,
                       double f1_left, double f1_right, double rho_ref) {
  double f1 = 0.0, theta_p = 1.0, theta_m = 1.0;
  int f1_computed = 0;
  const double up = alphaL * rhoYL - dt * fc;
  if (up < 0) {
    f1 = 0.5 * f1_left + 0.5 * f1_right;
    f1_computed = 1;
    const double up_lf = alphaL * rhoYL - dt * f1;
    if (fabs(up - up_lf) > 1e-20 * rho_ref) {
      theta_p = (0 - up_lf) / (up - up_lf);
      if (theta_p > 1)
        theta_p = 1.0;
      else if (theta_p < 0)
        theta_p = 0;
    }
  }
  const double um = alphaR * rhoYR + dt * fc;
  if (um < 0) {
    if (!f1_computed) f1 = 0.5 * f1_left + 0.5 * f1_right;
    const double um_lf = alphaR * rhoYR + dt * f1;
    if (fabs(um - um_lf) > 1e-20 * rho_ref) {
      theta_m = (0 - um_lf) / (um - um_lf);
      if (theta_m > 1)
        theta_m = 1.0;
      else if (theta_m < 0)
        theta_m = 0;
    }
  }
  return dmin(theta_p, theta_m) * (fc - f1) + f1;
}

/* Component-wise WENO (inviscid_scheme 51 / 71), src/WENO.cu:890-1086
 * (y 1619-1816, z 2284-2481): local Lax-Friedrichs splitting with the
 * point's own |U_k| + c|grad k| and WENO on each conservative flux, with the
 * species positivity limiter.  S[0..lo+hi] are the stencil points, S[lo] the
 * left node of the face. */
static void cp_weno_face(const oc_zone *z, const lay_t *L, const oc_run *r, const long *pts, int lo, int hi,
                         int dir, double eps_ref, int if_shock, double *fc, const int *ijk) {
  const int nv = r->n_var, npt = lo + hi + 1;
  const long sz = L->sz;
  double fp[OC_MAXV][8], fm[OC_MAXV][8];
  for (int m = 0; m < npt; ++m) {
    const long p = pts[m];
    const double q0 = z->cv[p], q1 = z->cv[p + sz], q2 = z->cv[p + 2 * sz], q3 = z->cv[p + 3 * sz],
                 q4 = z->cv[p + 4 * sz];
    double pk = z->bv[p + 4 * sz];
    const double m0 = z->metric[p + (3 * dir) * sz], m1 = z->metric[p + (3 * dir + 1) * sz],
                 m2 = z->metric[p + (3 * dir + 2) * sz];
    const double Uk = (q1 * m0 + q2 * m1 + q3 * m2) / q0;
    double cGradK = norm3(m0, m1, m2);
    if (r->mixture)
      cGradK *= z->acoustic[p];
    else
      cGradK *= sqrt(gamma_air * R_AIR * z->bv[p + 5 * sz]);
    const double lambda0 = fabs(Uk) + cGradK;
    const double halfJac = 0.5 * z->jac[p];
    const double sPlus = halfJac * (Uk + lambda0), sMinus = halfJac * (Uk - lambda0);
    pk *= halfJac;
    fp[0][m] = sPlus * q0;
    fp[1][m] = fma(sPlus, q1, pk * m0);
    fp[2][m] = fma(sPlus, q2, pk * m1);
    fp[3][m] = fma(sPlus, q3, pk * m2);
    fp[4][m] = fma(sPlus, q4, pk * Uk);
    fm[0][m] = sMinus * q0;
    fm[1][m] = fma(sMinus, q1, pk * m0);
    fm[2][m] = fma(sMinus, q2, pk * m1);
    fm[3][m] = fma(sMinus, q3, pk * m2);
    fm[4][m] = fma(sMinus, q4, pk * Uk);
    for (int l = 5; l < nv; ++l) {
      const double ql = z->cv[p + l * sz];
      fp[l][m] = sPlus * ql;
      fm[l][m] = sMinus * ql;
    }
  }
  const double vr2 = r->v_inf * r->v_inf;
  const double eps_scaled[3] = {eps_ref, eps_ref * vr2, eps_ref * vr2 * vr2};
  for (int l = 0; l < nv; ++l) {
    const double eps_here = (l >= 1 && l <= 3) ? eps_scaled[1] : (l == 4 ? eps_scaled[2] : eps_scaled[0]);
    double vp[7], vm[7];
    for (int m = 0; m < 2 * lo + 1; ++m) {
      vp[m] = fp[l][m];
      vm[m] = fm[l][m + 1];
    }
    fc[l] = lo == 3 ? oc_weno7(vp, vm, eps_here, if_shock) : oc_weno5(vp, vm, eps_here, if_shock);
  }
  if (r->positive_preserving && nv > 5) {
    const double alpha = z->mz == 1 ? 0.5 : 1.0 / 3.0;
    const long pl = pts[lo], pr = pts[lo + 1];
    const double jac_l = z->jac[pl], jac_r = z->jac[pr];
    const double alphaL = 0.5 * alpha * jac_l, alphaR = 0.5 * alpha * jac_r;
    const double dt = pp_time_step(z, r, ijk[0], ijk[1], ijk[2]);
    /* the first-order flux's advection speeds take the xi metric (components
     * 0..2) in every direction, src/WENO.cu:1017-1045, 1746-1774, 2411-2439 */
    const double temp1 = (z->metric[pl] * z->cv[pl + sz] + z->metric[pl + sz] * z->cv[pl + 2 * sz] +
                          z->metric[pl + 2 * sz] * z->cv[pl + 3 * sz]) / z->cv[pl] * jac_l;
    const double temp2 = (z->metric[pr] * z->cv[pr + sz] + z->metric[pr + sz] * z->cv[pr + 2 * sz] +
                          z->metric[pr + 2 * sz] * z->cv[pr + 3 * sz]) / z->cv[pr] * jac_r;
    for (int l = 0; l < nv - 5; ++l) {
      const double rhoYL = z->cv[pl + (l + 5) * sz], rhoYR = z->cv[pr + (l + 5) * sz];
      fc[l + 5] = pp_limit(fc[l + 5], alphaL, alphaR, rhoYL, rhoYR, dt, fp[5 + l][lo] + temp1 * rhoYL,
                           fm[5 + l][lo + 1] - temp2 * rhoYR, r->rho_inf);
    }
  }
}

/* Inviscid flux in one direction + its divergence into dq.
 * Flux kernels src/WENO.cu:831, 1561, 2225; differences 1532-1559, 2196-2223, 2859-2886.
 * Face metrics follow the reference: every direction averages the metric of
 * (i,j,k) and (i+1,j,k) (src/WENO.cu:886-888, 1615-1617, 2280-2282). */
static void weno_dir(oc_zone *z, const oc_run *r, const oc_tables *tb, int dir) {
  const lay_t L = layout(z);
  const int nv = r->n_var;
  const int n[3] = {z->mx, z->my, z->mz};
  const long stride[3] = {1, L.sx, L.sxy};
  double *flux = dir == 0 ? z->ff : (dir == 1 ? z->gf : z->hf);
  const int sch = r->inviscid_scheme;
  const int w7 = sch == 72 || sch == 71, charw = sch == 52 || sch == 72;
  const int lo = w7 ? 3 : 2, hi = w7 ? 4 : 3; /* stencil extent about the left point */
  const double eps_ref = eps_weno * r->weno_eps_scale;
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  /* faces */
#pragma omp parallel for schedule(static)
  for (int b = 0; b < n[d2]; ++b)
    for (int a = 0; a < n[d1]; ++a)
      for (int f = -1; f < n[dir]; ++f) {
        int idx[3];
        idx[dir] = f;
        idx[d1] = a;
        idx[d2] = b;
        const long p = IX(L, idx[0], idx[1], idx[2]);
        const int if_shock = face_if_shock(z, r, p, stride[dir]);
        double fc[OC_MAXV];
        if (!charw) {
          long pts[8];
          for (int m = -lo; m <= hi; ++m) pts[m + lo] = p + m * stride[dir];
          cp_weno_face(z, &L, r, pts, lo, hi, dir, eps_ref, if_shock, fc, idx);
        } else {
          spt_t S[8];
          for (int m = -lo; m <= hi; ++m) load_stencil_point(z, &L, r, p + m * stride[dir], dir, &S[m + lo]);
          const long pm = IX(L, idx[0] + 1, idx[1], idx[2]);
          const double kx = 0.5 * (z->metric[p + (3 * dir) * L.sz] + z->metric[pm + (3 * dir) * L.sz]);
          const double ky = 0.5 * (z->metric[p + (3 * dir + 1) * L.sz] + z->metric[pm + (3 * dir + 1) * L.sz]);
          const double kz = 0.5 * (z->metric[p + (3 * dir + 2) * L.sz] + z->metric[pm + (3 * dir + 2) * L.sz]);
          char_weno_face(r, tb, S, lo, kx, ky, kz, eps_ref, if_shock, fc);
          /* positivity limiter: x direction only (the y and z characteristic
           * kernels have none), src/WENO.cu:1459-1527 */
          if (dir == 0 && r->positive_preserving && r->mixture) {
            const spt_t *Lp = &S[lo], *Rp = &S[lo + 1];
            const double alpha = z->mz == 1 ? 0.5 : 1.0 / 3.0;
            const double alphaL = 0.5 * alpha * Lp->F[6], alphaR = 0.5 * alpha * Rp->F[6];
            const double dt = pp_time_step(z, r, idx[0], idx[1], idx[2]);
            const double U1 = Lp->F[0] * Lp->F[6], U2 = Rp->F[0] * Rp->F[6];
            for (int l = 0; l < nv - 5; ++l) {
              const double rhoYL = Lp->Q[l + 5], rhoYR = Rp->Q[l + 5];
              const double fpl = 0.5 * Lp->F[6] * (Lp->F[0] + Lp->F[5]) * rhoYL;
              const double fmr = 0.5 * Rp->F[6] * (Rp->F[0] - Rp->F[5]) * rhoYR;
              fc[l + 5] = pp_limit(fc[l + 5], alphaL, alphaR, rhoYL, rhoYR, dt, fpl + U1 * rhoYL, fmr - U2 * rhoYR,
                                   r->rho_inf);
            }
          }
        }
        for (int l = 0; l < nv; ++l) flux[p + l * L.sz] = fc[l];
      }
  /* divergence */
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k), q = DQX(z, i, j, k);
        for (int l = 0; l < nv; ++l)
          z->dq[q + l * dsz] -= flux[p + l * L.sz] - flux[p - stride[dir] + l * L.sz];
      }
}

/* Shock sensors, src/FieldOperation.cu:122-205, written over the interior by
 * compute_shock_sensor (207-228). */
static double ducros_sensor(const oc_zone *z, const lay_t *L, double eps, int i, int j, int k) {
  const long sz = L->sz, p = IX(*L, i, j, k);
  const double *m = z->metric;
  const double xi_x = m[p], xi_y = m[p + sz], xi_z = m[p + 2 * sz];
  const double eta_x = m[p + 3 * sz], eta_y = m[p + 4 * sz], eta_z = m[p + 5 * sz];
  const double zeta_x = m[p + 6 * sz], zeta_y = m[p + 7 * sz], zeta_z = m[p + 8 * sz];
  const double a = 1.0 / 12.0, b = 2.0 / 3.0;
  double d[3][3]; /* d[velocity component][computational direction] */
  const long st[3] = {1, L->sx, L->sxy};
  for (int c = 0; c < 3; ++c)
    for (int e = 0; e < 3; ++e) {
      const double *v = z->bv + (c + 1) * sz;
      d[c][e] = a * v[p - 2 * st[e]] - b * v[p - st[e]] + b * v[p + st[e]] - a * v[p + 2 * st[e]];
    }
  const double duDx = d[0][0] * xi_x + d[0][1] * eta_x + d[0][2] * zeta_x;
  const double dvDy = d[1][0] * xi_y + d[1][1] * eta_y + d[1][2] * zeta_y;
  const double dwDz = d[2][0] * xi_z + d[2][1] * eta_z + d[2][2] * zeta_z;
  double divV2 = duDx + dvDy + dwDz;
  if (divV2 > 0) return 0;
  const double duDy = d[0][0] * xi_y + d[0][1] * eta_y + d[0][2] * zeta_y;
  const double duDz = d[0][0] * xi_z + d[0][1] * eta_z + d[0][2] * zeta_z;
  const double dvDx = d[1][0] * xi_x + d[1][1] * eta_x + d[1][2] * zeta_x;
  const double dvDz = d[1][0] * xi_z + d[1][1] * eta_z + d[1][2] * zeta_z;
  const double dwDx = d[2][0] * xi_x + d[2][1] * eta_x + d[2][2] * zeta_x;
  const double dwDy = d[2][0] * xi_y + d[2][1] * eta_y + d[2][2] * zeta_y;
  divV2 *= divV2;
  const double curl2 = (dwDy - dvDz) * (dwDy - dvDz) + (duDz - dwDx) * (duDz - dwDx) + (dvDx - duDy) * (dvDx - duDy);
  return divV2 / (divV2 + curl2 + eps);
}

static double jameson_sensor(const oc_zone *z, const lay_t *L, int i, int j, int k) {
  const long p = IX(*L, i, j, k);
  const double *pr = z->bv + 4 * L->sz;
  const double twoP = pr[p] * 2;
  const double pI = pr[p + 1] + pr[p - 1], pJ = pr[p + L->sx] + pr[p - L->sx], pK = pr[p + L->sxy] + pr[p - L->sxy];
  return fabs(pI - twoP) / (pI + twoP) + fabs(pJ - twoP) / (pJ + twoP) + fabs(pK - twoP) / (pK + twoP);
}

static double jump_sensor(const oc_zone *z, const lay_t *L, int i, int j, int k) {
  const long p = IX(*L, i, j, k);
  const double *rho = z->bv, *pr = z->bv + 4 * L->sz;
  const double rhoI = 1.0 / rho[p], pI = 1.0 / pr[p];
  double s = 0;
  s = dmax(s, fabs(rho[p + 1] - rho[p]) * rhoI);
  s = dmax(s, fabs(rho[p + L->sx] - rho[p]) * rhoI);
  s = dmax(s, fabs(rho[p + L->sxy] - rho[p]) * rhoI);
  s = dmax(s, fabs(pr[p + 1] - pr[p]) * pI);
  s = dmax(s, fabs(pr[p + L->sx] - pr[p]) * pI);
  s = dmax(s, fabs(pr[p + L->sxy] - pr[p]) * pI);
  return s;
}

void oc_shock_sensor(oc_zone *z, const oc_run *r) {
  const lay_t L = layout(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        double s;
        if (r->shock_sensor == 1)
          s = jameson_sensor(z, &L, i, j, k);
        else if (r->shock_sensor == 2)
          s = jump_sensor(z, &L, i, j, k);
        else
          s = ducros_sensor(z, &L, r->sensor_eps, i, j, k);
        z->shock_sensor[IX(L, i, j, k)] = s;
      }
}

/* ------------------------------------------------------------------------ */
/* primitive-variable path: reconstruction + LF / AUSM+ / HLLC, and Roe      */
/* compute_convective_term_pv_1D, src/InviscidScheme.cu:61-169;             */
/* Roe_compute_inviscid_flux_1D 306-395; solvers 397-590, 626-787            */
/* ------------------------------------------------------------------------ */
static double sgn_real(double a) { return a < 0 ? -1.0 : 1.0; } /* gxl::sgn */

/* Limiter<0/1>::apply_limiter, src/Limiter.cuh:14-16 (minmod for both) */
static double minmod(double a, double b) { return 0.5 * (sgn_real(a) + sgn_real(b)) * dmin(fabs(a), fabs(b)); }

/* point data of the primitive path: rho, u, v, w, p, Y (nv values) */
typedef struct {
  double q[OC_MAXV];
  double m[3]; /* metric of the sweep direction */
  double jac;
  double efd; /* entropy-fix delta (Roe) */
} ppt_t;

/* compute_mixture_characteristic_thermo outputs e_tr and gamma, src/Thermo.cuh:49-87 */
static void mixture_e_gamma(const oc_tables *tb, double t, const double *y, double *e_tr, double *gamma,
                            double *c) {
  const double t_eval = dmax(t, 1.0);
  double h[OC_MAX_SPEC], cp[OC_MAX_SPEC];
  oc_enthalpy_and_cp(tb, t_eval, h, cp);
  double e = 0, rl = 0, cpl = 0, cvl = 0;
  for (int l = 0; l < tb->n_spec; ++l) {
    const double h_tr = h[l] - oc_ve_energy(tb, l, t_eval), cp_tr = cp[l] - oc_ve_cv(tb, l, t_eval);
    e += y[l] * (h_tr - tb->gas_const[l] * t_eval);
    rl += y[l] * tb->gas_const[l];
    cpl += y[l] * cp_tr;
    cvl += y[l] * (cp_tr - tb->gas_const[l]);
  }
  const double g = cpl / dmax(cvl, 1e-8);
  if (e_tr) *e_tr = e;
  if (gamma) *gamma = g;
  if (c) *c = sqrt(dmax(g * rl * t_eval, 1e-12));
}

/* reconstruction<mix>, src/InviscidScheme.cu:171-219 and src/Reconstruction.cu:5-44:
 * left/right states at the face between window points 1 and 2 of P[0..3],
 * then rho E at index nv and gamma at nv+1. */
static void reconstruct(const oc_run *r, const oc_tables *tb, const ppt_t *P, double *pl, double *pr) {
  const int nv = r->n_var;
  for (int l = 0; l < nv; ++l) {
    const double a = P[0].q[l], b = P[1].q[l], c = P[2].q[l], d = P[3].q[l];
    if (r->reconstruction == 2) {
      const double kappa = 1.0 / 3.0;
      const double di = b - a, di1 = c - b, di2 = d - c;
      const double neg_l = minmod(di, di1), pos_l = minmod(di1, di);
      const double neg_r = minmod(di1, di2), pos_r = minmod(di2, di1);
      pl[l] = b + 0.25 * ((1 - kappa) * neg_l + (1 + kappa) * pos_l);
      pr[l] = c - 0.25 * ((1 - kappa) * pos_r + (1 + kappa) * neg_r);
    } else if (r->reconstruction == 3) {
      const double di = b - a, di1 = c - b, di2 = d - c;
      pl[l] = b + 0.5 * minmod(di, di1);
      pr[l] = c - 0.5 * minmod(di1, di2);
    } else {
      pl[l] = b;
      pr[l] = c;
    }
  }
  if (r->mixture) {
    const double ek_l = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    const double ek_r = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    double rl = 0, rr = 0;
    for (int l = 0; l < r->n_spec; ++l) {
      rl += pl[5 + l] * tb->gas_const[l];
      rr += pr[5 + l] * tb->gas_const[l];
    }
    const double t_l = pl[4] / dmax(pl[0] * rl, 1e-12), t_r = pr[4] / dmax(pr[0] * rr, 1e-12);
    double e_l, e_r, g_l, g_r;
    mixture_e_gamma(tb, t_l, pl + 5, &e_l, &g_l, NULL);
    mixture_e_gamma(tb, t_r, pr + 5, &e_r, &g_r, NULL);
    const double eve_l = r->two_temperature ? dmax(pl[5 + r->n_spec], 0.0) : 0.0;
    const double eve_r = r->two_temperature ? dmax(pr[5 + r->n_spec], 0.0) : 0.0;
    pl[nv] = pl[0] * (ek_l + e_l + eve_l);
    pr[nv] = pr[0] * (ek_r + e_r + eve_r);
    pl[nv + 1] = g_l;
    pr[nv + 1] = g_r;
  } else {
    const double el = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    const double er = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    pl[nv] = el * pl[0] + pl[4] / (gamma_air - 1);
    pr[nv] = er * pr[0] + pr[4] / (gamma_air - 1);
  }
}

/* WENO5-JS / WENO7-Z reconstruction of the primitive variables for the
 * AUSM+ / HLLC / Roe solvers (reconstruction 4 / 5; no reference counterpart:
 * the reference's WENO is flux-split).  C points at the face's left node,
 * C[-w+1 .. w] are valid.  Each half is the matching half of oc_weno5 /
 * oc_weno7 (the other half fed zeros adds exactly 0).  epsilon is 1e-40 times
 * the squared scale of the variable (rho_ref, v_ref, rho_ref v_ref^2, 1 for
 * scalars).  A state with a non-positive density or pressure falls back to
 * the first-order states.  rho E and gamma follow as in reconstruct(). */
static void weno_reconstruct(const oc_run *r, const oc_tables *tb, const ppt_t *C, double *pl, double *pr) {
  const int nv = r->n_var, w7 = r->reconstruction == 5;
  const double vr2 = r->v_inf * r->v_inf;
  const double zero[7] = {0, 0, 0, 0, 0, 0, 0};
  int ok = 1;
  for (int l = 0; l < nv; ++l) {
    const double eps = l == 0 ? 1e-40 * r->rho_inf * r->rho_inf
                              : (l < 4 ? 1e-40 * vr2 : (l == 4 ? 1e-40 * (r->rho_inf * vr2) * (r->rho_inf * vr2) : 1e-40));
    double v[7], m[7];
    if (w7) {
      for (int k = 0; k < 7; ++k) {
        v[k] = C[k - 3].q[l];
        m[k] = C[k - 2].q[l];
      }
      pl[l] = oc_weno7(v, zero, eps, 1);
      pr[l] = oc_weno7(zero, m, eps, 1);
    } else {
      for (int k = 0; k < 5; ++k) {
        v[k] = C[k - 2].q[l];
        m[k] = C[k - 1].q[l];
      }
      pl[l] = oc_weno5(v, zero, eps, 1);
      pr[l] = oc_weno5(zero, m, eps, 1);
    }
    if ((l == 0 || l == 4) && !(pl[l] > 0 && pr[l] > 0)) ok = 0;
  }
  if (!ok)
    for (int l = 0; l < nv; ++l) {
      pl[l] = C[0].q[l];
      pr[l] = C[1].q[l];
    }
  /* energy and gamma of the states: the post-processing of reconstruct() */
  oc_run r1 = *r;
  r1.reconstruction = 1;
  ppt_t W[4];
  for (int l = 0; l < nv; ++l) {
    W[0].q[l] = W[1].q[l] = pl[l];
    W[2].q[l] = W[3].q[l] = pr[l];
  }
  reconstruct(&r1, tb, W, pl, pr);
}

/* riemannSolver_laxFriedrich<Air>, src/InviscidScheme.cu:753-787 */
static void solver_lf(const oc_run *r, const double *pl, const double *pr, const ppt_t *A, const ppt_t *B,
                      double *fc) {
  const int nv = r->n_var;
  const double kx = 0.5 * (A->m[0] + B->m[0]), ky = 0.5 * (A->m[1] + B->m[1]), kz = 0.5 * (A->m[2] + B->m[2]);
  const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ukl = pl[1] * kx + pl[2] * ky + pl[3] * kz, Ukr = pr[1] * kx + pr[2] * ky + pr[3] * kz;
  const double cl = sqrt(gamma_air * pl[4] / pl[0]), cr = sqrt(gamma_air * pr[4] / pr[0]);
  const double sr = dmax(fabs(Ukl) + cl * gradK, fabs(Ukr) + cr * gradK);
  const double hj = 0.5 * 0.5 * (A->jac + B->jac);
  const double rul = pl[0] * Ukl, rur = pr[0] * Ukr;
  fc[0] = (rul + rur - sr * (pr[0] - pl[0])) * hj;
  fc[1] = (rul * pl[1] + rur * pr[1] + kx * (pl[4] + pr[4]) - sr * (pr[1] * pr[0] - pl[1] * pl[0])) * hj;
  fc[2] = (rul * pl[2] + rur * pr[2] + ky * (pl[4] + pr[4]) - sr * (pr[2] * pr[0] - pl[2] * pl[0])) * hj;
  fc[3] = (rul * pl[3] + rur * pr[3] + kz * (pl[4] + pr[4]) - sr * (pr[3] * pr[0] - pl[3] * pl[0])) * hj;
  fc[4] = ((pl[nv] + pl[4]) * Ukl + (pr[nv] + pr[4]) * Ukr - sr * (pr[nv] - pl[nv])) * hj;
}

/* riemannSolver_ausmPlus, src/InviscidScheme.cu:397-469 */
static void solver_ausm(const oc_run *r, const double *pl, const double *pr, const ppt_t *A, const ppt_t *B,
                        double *fc) {
  const int nv = r->n_var;
  const double k1 = 0.5 * (A->jac * A->m[0] + B->jac * B->m[0]);
  const double k2 = 0.5 * (A->jac * A->m[1] + B->jac * B->m[1]);
  const double k3 = 0.5 * (A->jac * A->m[2] + B->jac * B->m[2]);
  const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk, ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
  const double p_l = pl[4], p_r = pr[4], rho_l = pl[0], rho_r = pr[0];
  const double gam_l = r->mixture ? pl[nv + 1] : gamma_air, gam_r = r->mixture ? pr[nv + 1] : gamma_air;
  const double c = 0.5 * (sqrt(gam_l * p_l / rho_l) + sqrt(gam_r * p_r / rho_r));
  const double ml = ul / c, mr = ur / c;
  const double alpha = 3.0 / 16.0;
  double mlp, plp, mrn, prn;
  if (fabs(ml) > 1) {
    mlp = 0.5 * (ml + fabs(ml));
    plp = mlp / ml;
  } else {
    const double a = (ml + 1) * (ml + 1) * 0.25, b = (ml * ml - 1) * (ml * ml - 1);
    mlp = a + 0.125 * b;
    plp = a * (2 - ml) + alpha * ml * b;
  }
  if (fabs(mr) > 1) {
    mrn = 0.5 * (mr - fabs(mr));
    prn = mrn / mr;
  } else {
    const double a = (mr - 1) * (mr - 1) * 0.25, b = (mr * mr - 1) * (mr * mr - 1);
    mrn = -a - 0.125 * b;
    prn = a * (2 + mr) - alpha * mr * b;
  }
  const double p_coeff = plp * p_l + prn * p_r;
  const double m_half = mlp + mrn;
  const double mpos = 0.5 * (m_half + fabs(m_half)), mneg = 0.5 * (m_half - fabs(m_half));
  const double mass = c * (rho_l * mpos + rho_r * mneg);
  const double coeff = mass * gk;
  const double *s = mass >= 0 ? pl : pr;
  fc[0] = coeff;
  fc[1] = coeff * s[1] + p_coeff * k1;
  fc[2] = coeff * s[2] + p_coeff * k2;
  fc[3] = coeff * s[3] + p_coeff * k3;
  fc[4] = coeff * (s[nv] + s[4]) / s[0];
  for (int l = 5; l < nv; ++l) fc[l] = coeff * s[l];
}

/* riemannSolver_hllc, src/InviscidScheme.cu:471-590 */
static void solver_hllc(const oc_run *r, const oc_tables *tb, const double *pl, const double *pr, const ppt_t *A,
                        const ppt_t *B, double *fc) {
  const int nv = r->n_var;
  const double sl0 = sqrt(pl[0]), sr0 = sqrt(pr[0]);
  const double rl_c = sl0 / (sl0 + sr0), rr_c = sr0 / (sl0 + sr0);
  const double ut = pl[1] * rl_c + pr[1] * rr_c, vt = pl[2] * rl_c + pr[2] * rr_c, wt = pl[3] * rl_c + pr[3] * rr_c;
  double c_t;
  if (!r->mixture) {
    const double hl = (pl[nv] + pl[4]) / pl[0], hr = (pr[nv] + pr[4]) / pr[0];
    const double ht = hl * rl_c + hr * rr_c;
    const double ekt = 0.5 * (ut * ut + vt * vt + wt * wt);
    c_t = sqrt((gamma_air - 1) * (ht - ekt));
  } else {
    double svm[OC_MAXV] = {0};
    for (int l = 0; l < nv - 5; ++l) svm[l] = rl_c * pl[l + 5] + rr_c * pr[l + 5];
    double rg = 0;
    for (int l = 0; l < r->n_spec; ++l) rg += svm[l] * tb->gas_const[l];
    const double tl = pl[4] / pl[0], tr = pr[4] / pr[0];
    const double t = (rl_c * tl + rr_c * tr) / dmax(rg, 1e-12);
    double g;
    mixture_e_gamma(tb, t, svm, NULL, &g, &c_t);
  }
  const double kx = 0.5 * (A->m[0] + B->m[0]), ky = 0.5 * (A->m[1] + B->m[1]), kz = 0.5 * (A->m[2] + B->m[2]);
  const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ut = (kx * ut + ky * vt + kz * wt) / gradK;
  const double jac_ave = 0.5 * (A->jac + B->jac);
  const double g_l = r->mixture ? pl[nv + 1] : gamma_air, g_r = r->mixture ? pr[nv + 1] : gamma_air;
  const double Ul = kx * pl[1] + ky * pl[2] + kz * pl[3];
  const double cl = sqrt(g_l * pl[4] / pl[0]);
  const double sl = dmin(Ul / gradK - cl, Ut - c_t);
  if (sl >= 0) {
    const double ru = pl[0] * Ul;
    fc[0] = jac_ave * ru;
    fc[1] = jac_ave * (ru * pl[1] + pl[4] * kx);
    fc[2] = jac_ave * (ru * pl[2] + pl[4] * ky);
    fc[3] = jac_ave * (ru * pl[3] + pl[4] * kz);
    fc[4] = jac_ave * ((pl[nv] + pl[4]) * Ul);
    for (int l = 5; l < nv; ++l) fc[l] = jac_ave * ru * pl[l];
    return;
  }
  const double Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
  const double cr = sqrt(g_r * pr[4] / pr[0]);
  const double sr = dmax(Ur / gradK + cr, Ut + c_t);
  if (sr < 0) {
    const double ru = pr[0] * Ur;
    fc[0] = jac_ave * ru;
    fc[1] = jac_ave * (ru * pr[1] + pr[4] * kx);
    fc[2] = jac_ave * (ru * pr[2] + pr[4] * ky);
    fc[3] = jac_ave * (ru * pr[3] + pr[4] * kz);
    fc[4] = jac_ave * ((pr[nv] + pr[4]) * Ur);
    for (int l = 5; l < nv; ++l) fc[l] = jac_ave * ru * pr[l];
    return;
  }
  const double sm = ((pr[0] * Ur * (sr - Ur / gradK) - pl[0] * Ul * (sl - Ul / gradK)) / gradK + pl[4] - pr[4]) /
                    (pr[0] * (sr - Ur / gradK) - pl[0] * (sl - Ul / gradK));
  const double pm = pl[0] * (sl - Ul / gradK) * (sm - Ul / gradK) + pl[4];
  const double *q = sm >= 0 ? pl : pr;
  const double s = sm >= 0 ? sl : sr, U = sm >= 0 ? Ul : Ur;
  const double pc = 1.0 / (s - sm);
  const double Qc = jac_ave * pc * sm * (s * gradK - U) * q[0];
  fc[0] = Qc;
  const double dP = (s * pm - sm * q[4]) * pc * jac_ave;
  fc[1] = Qc * q[1] + dP * kx;
  fc[2] = Qc * q[2] + dP * ky;
  fc[3] = Qc * q[3] + dP * kz;
  fc[4] = Qc * q[nv] / q[0] + pc * jac_ave * (s * pm * sm * gradK - sm * q[4] * U);
  for (int l = 5; l < nv; ++l) fc[l] = Qc * q[l];
}

/* riemannSolver_Roe with compute_half_sum_left_right_flux,
 * src/InviscidScheme.cu:592-746 */
static void solver_roe(const oc_run *r, const oc_tables *tb, const double *pl, const double *pr, const ppt_t *A,
                       const ppt_t *B, double *fc) {
  const int nv = r->n_var;
  /* half sum of the left and right physical fluxes (own metric and jacobian) */
  const ppt_t *S[2] = {A, B};
  const double *Q[2] = {pl, pr};
  for (int l = 0; l < nv; ++l) fc[l] = 0;
  for (int side = 0; side < 2; ++side) {
    const double *q = Q[side];
    const double jx = S[side]->jac * S[side]->m[0], jy = S[side]->jac * S[side]->m[1],
                 jz = S[side]->jac * S[side]->m[2];
    const double Uk = q[1] * jx + q[2] * jy + q[3] * jz;
    const double coeff = Uk * q[0];
    fc[0] += 0.5 * coeff;
    fc[1] += 0.5 * (coeff * q[1] + q[4] * jx);
    fc[2] += 0.5 * (coeff * q[2] + q[4] * jy);
    fc[3] += 0.5 * (coeff * q[3] + q[4] * jz);
    fc[4] += 0.5 * Uk * (q[4] + q[nv]);
    for (int l = 5; l < nv; ++l) fc[l] += 0.5 * coeff * q[l];
  }
  const double dl = sqrt(pl[0]), dr = sqrt(pr[0]);
  const double inv = 1.0 / (dl + dr);
  const double u = (dl * pl[1] + dr * pr[1]) * inv, v = (dl * pl[2] + dr * pr[2]) * inv,
               w = (dl * pl[3] + dr * pr[3]) * inv;
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double hl = (pl[nv] + pl[4]) / pl[0], hr = (pr[nv] + pr[4]) / pr[0];
  const double h = (dl * hl + dr * hr) * inv;
  double gamma = gamma_air;
  double c = sqrt((gamma - 1) * (h - ek));
  double svm[OC_MAXV] = {0}, alpha[OC_MAXV] = {0}, ecoef[OC_MAXV] = {0};
  for (int l = 0; l < nv - 5; ++l) svm[l] = (dl * pl[l + 5] + dr * pr[l + 5]) * inv;
  if (r->mixture) {
    double rg = 0;
    for (int l = 0; l < r->n_spec; ++l) rg += svm[l] * tb->gas_const[l];
    const double tl = pl[4] / pl[0], tr = pr[4] / pr[0];
    const double t = (dl * tl + dr * tr) * inv / dmax(rg, 1e-12);
    mixture_char_thermo(tb, r, t, svm, alpha, ecoef, &gamma, &c);
  }
  double kx = 0.5 * (A->m[0] + B->m[0]), ky = 0.5 * (A->m[1] + B->m[1]), kz = 0.5 * (A->m[2] + B->m[2]);
  const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
  double Uk = kx * u + ky * v + kz * w;
  double ch[3] = {Uk - gradK * c, Uk, Uk + gradK * c};
  const double efd = 0.5 * (A->efd + B->efd);
  for (int m = 0; m < 3; ++m) {
    ch[m] = fabs(ch[m]);
    if (ch[m] < efd) ch[m] = 0.5 * (ch[m] * ch[m] / efd + efd);
  }
  kx /= gradK;
  ky /= gradK;
  kz /= gradK;
  Uk /= gradK;
  const double jac_ave = 0.5 * (A->jac + B->jac);
  double dq[OC_MAXV + 1] = {0};
  dq[0] = jac_ave * (pr[0] - pl[0]);
  for (int l = 1; l < nv; ++l) dq[l] = jac_ave * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = jac_ave * (pr[nv] - pl[nv]);
  double c1 = (gamma - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
  double c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - Uk * dq[0]) / c;
  for (int l = 0; l < r->n_scalar_transported; ++l) c1 += alpha[l] * dq[l + 5];
  double c3 = dq[0] - c1;
  double LDq[OC_MAXV] = {0};
  LDq[0] = 0.5 * (c1 - c2);
  LDq[1] = kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c;
  LDq[2] = ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c;
  LDq[3] = kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c;
  LDq[4] = 0.5 * (c1 + c2);
  for (int l = 0; l < r->n_scalar_transported; ++l) LDq[l + 5] = dq[l + 5] - svm[l] * dq[0];
  double b[OC_MAXV];
  b[0] = -ch[0] * LDq[0];
  for (int l = 1; l < nv; ++l) b[l] = -ch[1] * LDq[l];
  b[4] = -ch[2] * LDq[4];
  const double c0 = kx * b[1] + ky * b[2] + kz * b[3];
  c1 = c0 + b[0] + b[4];
  c2 = c * (b[4] - b[0]);
  c3 = 0;
  for (int l = 0; l < r->n_scalar_transported; ++l) c3 += ecoef[l] * b[l + 5];
  fc[0] += 0.5 * c1;
  fc[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * b[2] - ky * b[3]));
  fc[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * b[3] - kz * b[1]));
  fc[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * b[1] - kx * b[2]));
  fc[4] += 0.5 * (h * c1 + Uk * c2 - c * c * c0 / (gamma - 1) +
                  c * ((kz * v - ky * w) * b[1] + (kx * w - kz * u) * b[2] + (ky * u - kx * v) * b[3]) + c3);
  for (int l = 0; l < nv - 5; ++l) fc[5 + l] += 0.5 * (b[l + 5] + svm[l] * c1);
}

/* compute_entropy_fix_delta, src/InviscidScheme.cu:269-304, at one node */
static double entropy_fix_delta(const oc_zone *z, const lay_t *L, const oc_run *r, long p) {
  const long sz = L->sz;
  const double *mt = z->metric, *bv = z->bv;
  const double u = bv[p + sz], v = bv[p + 2 * sz], w = bv[p + 3 * sz];
  const double U = fabs(u * mt[p] + v * mt[p + sz] + w * mt[p + 2 * sz]);
  const double V = fabs(u * mt[p + 3 * sz] + v * mt[p + 4 * sz] + w * mt[p + 5 * sz]);
  const double W = fabs(u * mt[p + 6 * sz] + v * mt[p + 7 * sz] + w * mt[p + 8 * sz]);
  const double kx = norm3(mt[p], mt[p + sz], mt[p + 2 * sz]);
  const double ky = norm3(mt[p + 3 * sz], mt[p + 4 * sz], mt[p + 5 * sz]);
  const double kz = norm3(mt[p + 6 * sz], mt[p + 7 * sz], mt[p + 8 * sz]);
  const double ac = r->mixture ? z->acoustic[p] : sqrt(gamma_air * bv[p + 4 * sz] / bv[p]);
  if (z->mz == 1) return r->entropy_fix_factor * (U + V + ac * 0.5 * (kx + ky));
  return r->entropy_fix_factor * (U + V + W + ac * (kx + ky + kz) / 3.0);
}

static void pv_dir(oc_zone *z, const oc_run *r, const oc_tables *tb, int dir) {
  const lay_t L = layout(z);
  const int nv = r->n_var;
  const int n[3] = {z->mx, z->my, z->mz};
  const long stride[3] = {1, L.sx, L.sxy};
  double *flux = dir == 0 ? z->ff : (dir == 1 ? z->gf : z->hf);
  const int d1 = dir == 0 ? 1 : 0, d2 = dir == 2 ? 1 : 2;
  const int roe = r->inviscid_scheme == 2;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < n[d2]; ++b)
    for (int a = 0; a < n[d1]; ++a)
      for (int f = -1; f < n[dir]; ++f) {
        int idx[3];
        idx[dir] = f;
        idx[d1] = a;
        idx[d2] = b;
        const long p = IX(L, idx[0], idx[1], idx[2]);
        /* window nodes -w+1 .. w about the face (P[3] its left node) */
        const int w = r->reconstruction == 5 ? 4 : (r->reconstruction == 4 ? 3 : 2);
        ppt_t P[8];
        for (int m = 4 - w; m < 4 + w; ++m) {
          const long q = p + (m - 3) * stride[dir];
          for (int l = 0; l < 5; ++l) P[m].q[l] = z->bv[q + l * L.sz];
          for (int l = 0; l < r->n_scalar; ++l) P[m].q[5 + l] = z->sv[q + l * L.sz];
          for (int l = 0; l < 3; ++l) P[m].m[l] = z->metric[q + (3 * dir + l) * L.sz];
          P[m].jac = z->jac[q];
          P[m].efd = (roe && (m == 3 || m == 4)) ? entropy_fix_delta(z, &L, r, q) : 0.0;
        }
        double pl[OC_MAXV + 2], pr[OC_MAXV + 2], fc[OC_MAXV];
        if (r->reconstruction >= 4)
          weno_reconstruct(r, tb, P + 3, pl, pr);
        else
          reconstruct(r, tb, P + 2, pl, pr);
        switch (r->inviscid_scheme) {
          case 1: solver_lf(r, pl, pr, &P[3], &P[4], fc); break;
          case 2: solver_roe(r, tb, pl, pr, &P[3], &P[4], fc); break;
          case 4: solver_hllc(r, tb, pl, pr, &P[3], &P[4], fc); break;
          default: solver_ausm(r, pl, pr, &P[3], &P[4], fc); break;
        }
        for (int l = 0; l < nv; ++l) flux[p + l * L.sz] = fc[l];
      }
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k), q = DQX(z, i, j, k);
        for (int l = 0; l < nv; ++l)
          z->dq[q + l * dsz] -= flux[p + l * L.sz] - flux[p - stride[dir] + l * L.sz];
      }
}

/* compute_inviscid_flux, src/InviscidScheme.cu:12-26: component-wise (51/71)
 * or characteristic (52/72) WENO or the primitive-variable path (1 LF, 3 AUSM+,
 * 4 HLLC) or Roe (2), one direction at a time. */
void oc_inviscid_dir(oc_zone *z, const oc_run *r, const oc_tables *tb, int dir) {
  const int sch = r->inviscid_scheme;
  if (sch == 51 || sch == 52 || sch == 71 || sch == 72)
    weno_dir(z, r, tb, dir);
  else
    pv_dir(z, r, tb, dir);
}

/* ------------------------------------------------------------------------ */
/* 8th-order collocated viscous fluxes, src/ViscousScheme.cu:140-834         */
/* ------------------------------------------------------------------------ */
/* d_dXi/d_dEta/d_dZeta<8>, src/ViscousScheme.cu:140-188 */
static double d8(const double *f, long p, long s, int i, int n, int bl, int br) {
#define F_(o) f[p + (long)(o) * s]
  double df = 0;
  int computed = 0;
  if (bl) {
    if (i == 0) {
      df = -25.0 / 12 * F_(0) + 4.0 * F_(1) - 3 * F_(2) + 4.0 / 3 * F_(3) - 0.25 * F_(4);
      computed = 1;
    } else if (i == 1) {
      df = -0.25 * F_(-1) - 5.0 / 6 * F_(0) + 1.5 * F_(1) - 0.5 * F_(2) + 1.0 / 12 * F_(3);
      computed = 1;
    } else if (i == 2) {
      df = -1.0 / 12 * (F_(2) - F_(-2)) + 2.0 / 3 * (F_(1) - F_(-1));
      computed = 1;
    } else if (i == 3) {
      df = 0.75 * (F_(1) - F_(-1)) - 0.15 * (F_(2) - F_(-2)) + 1.0 / 60 * (F_(3) - F_(-3));
      computed = 1;
    }
  }
  if (br) {
    if (i == n - 1) {
      df = 25.0 / 12 * F_(0) - 4.0 * F_(-1) + 3 * F_(-2) - 4.0 / 3 * F_(-3) + 0.25 * F_(-4);
      computed = 1;
    } else if (i == n - 2) {
      df = 0.25 * F_(1) + 5.0 / 6 * F_(0) - 1.5 * F_(-1) + 0.5 * F_(-2) - 1.0 / 12 * F_(-3);
      computed = 1;
    } else if (i == n - 3) {
      df = -1.0 / 12 * (F_(2) - F_(-2)) + 2.0 / 3 * (F_(1) - F_(-1));
      computed = 1;
    } else if (i == n - 4) {
      df = 0.75 * (F_(1) - F_(-1)) - 0.15 * (F_(2) - F_(-2)) + 1.0 / 60 * (F_(3) - F_(-3));
      computed = 1;
    }
  }
  if (!computed) {
    const double a1 = 0.8, a2 = -0.2, a3 = 4.0 / 105, a4 = -1.0 / 280;
    df = a1 * (F_(1) - F_(-1)) + a2 * (F_(2) - F_(-2)) + a3 * (F_(3) - F_(-3)) + a4 * (F_(4) - F_(-4));
  }
  return df;
#undef F_
}

static void btype_flags(const oc_zone *z, int i, int j, int k, int *ct) {
  const int mx = z->mx, my = z->my;
  ct[0] = z->btype_il[j + k * my] != 0;
  ct[1] = z->btype_ir[j + k * my] != 0;
  ct[2] = z->btype_jl[i + k * mx] != 0;
  ct[3] = z->btype_jr[i + k * mx] != 0;
  ct[4] = z->btype_kl[i + j * mx] != 0;
  ct[5] = z->btype_kr[i + j * mx] != 0;
}

/* periodic_exchange(task 1), src/DataCommunication.cu:665-703 */
static void periodic_exchange_flux(oc_zone *z, const oc_run *r, const oc_boundary *b) {
  const lay_t L = layout(z);
  const int g = z->ngg;
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  double *arrs[3] = {z->ff, z->gf, z->hf};
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        int o[3] = {i, j, k};
        const int nn[3] = {z->mx, z->my, z->mz};
        o[b->face] = b->direction < 0 ? nn[b->face] - 1 : 0;
        const long p = IX(L, i, j, k), po = IX(L, o[0], o[1], o[2]);
        for (int l = 0; l < r->n_var; ++l)
          for (int a = 0; a < 3; ++a) {
            double *f = arrs[a];
            const double ave = 0.5 * (f[p + l * L.sz] + f[po + l * L.sz]);
            f[p + l * L.sz] = ave;
            f[po + l * L.sz] = ave;
          }
        for (int gg = 1; gg <= g; ++gg) {
          const long pg = IX(L, i + gg * dir[0], j + gg * dir[1], k + gg * dir[2]);
          const long pi = IX(L, o[0] + gg * dir[0], o[1] + gg * dir[1], o[2] + gg * dir[2]);
          for (int l = 0; l < r->n_var; ++l)
            for (int a = 0; a < 3; ++a) arrs[a][pg + l * L.sz] = arrs[a][pi + l * L.sz];
        }
      }
}

static void viscous8_flux(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  const lay_t L = layout(z);
  const int mx = z->mx, my = z->my, mz = z->mz, ns = r->n_spec;
  const long sz = L.sz, si = 1, sj = L.sx, sk = L.sxy;
  /* compute_viscous_flux_collocated<mix,8>, src/ViscousScheme.cu:470-585 */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        int ct[6];
        btype_flags(z, i, j, k, ct);
        const long p = IX(L, i, j, k);
        const double *pv = z->bv;
#define DXI(arr, l) d8(arr, p + (long)(l) * sz, si, i, mx, ct[0], ct[1])
#define DETA(arr, l) d8(arr, p + (long)(l) * sz, sj, j, my, ct[2], ct[3])
#define DZETA(arr, l) d8(arr, p + (long)(l) * sz, sk, k, mz, ct[4], ct[5])
        const double u_xi = DXI(pv, 1), u_eta = DETA(pv, 1), u_zeta = DZETA(pv, 1);
        const double v_xi = DXI(pv, 2), v_eta = DETA(pv, 2), v_zeta = DZETA(pv, 2);
        const double w_xi = DXI(pv, 3), w_eta = DETA(pv, 3), w_zeta = DZETA(pv, 3);
        const double t_xi = DXI(pv, 5), t_eta = DETA(pv, 5), t_zeta = DZETA(pv, 5);
        const double *mt = z->metric;
        const double xi_x = mt[p], xi_y = mt[p + sz], xi_z = mt[p + 2 * sz];
        const double eta_x = mt[p + 3 * sz], eta_y = mt[p + 4 * sz], eta_z = mt[p + 5 * sz];
        const double zeta_x = mt[p + 6 * sz], zeta_y = mt[p + 7 * sz], zeta_z = mt[p + 8 * sz];
        const double u_x = u_xi * xi_x + u_eta * eta_x + u_zeta * zeta_x;
        const double u_y = u_xi * xi_y + u_eta * eta_y + u_zeta * zeta_y;
        const double u_z = u_xi * xi_z + u_eta * eta_z + u_zeta * zeta_z;
        const double v_x = v_xi * xi_x + v_eta * eta_x + v_zeta * zeta_x;
        const double v_y = v_xi * xi_y + v_eta * eta_y + v_zeta * zeta_y;
        const double v_z = v_xi * xi_z + v_eta * eta_z + v_zeta * zeta_z;
        const double w_x = w_xi * xi_x + w_eta * eta_x + w_zeta * zeta_x;
        const double w_y = w_xi * xi_y + w_eta * eta_y + w_zeta * zeta_y;
        const double w_z = w_xi * xi_z + w_eta * eta_z + w_zeta * zeta_z;
        const double mul = z->mul[p];
        const double tau_xx = mul * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
        const double tau_yy = mul * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
        const double tau_zz = mul * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
        const double tau_xy = mul * (u_y + v_x);
        const double tau_xz = mul * (u_z + w_x);
        const double tau_yz = mul * (v_z + w_y);
        double *fv = z->ff, *gv = z->gf, *hv = z->hf;
        fv[p + sz] = xi_x * tau_xx + xi_y * tau_xy + xi_z * tau_xz;
        fv[p + 2 * sz] = xi_x * tau_xy + xi_y * tau_yy + xi_z * tau_yz;
        fv[p + 3 * sz] = xi_x * tau_xz + xi_y * tau_yz + xi_z * tau_zz;
        gv[p + sz] = eta_x * tau_xx + eta_y * tau_xy + eta_z * tau_xz;
        gv[p + 2 * sz] = eta_x * tau_xy + eta_y * tau_yy + eta_z * tau_yz;
        gv[p + 3 * sz] = eta_x * tau_xz + eta_y * tau_yz + eta_z * tau_zz;
        hv[p + sz] = zeta_x * tau_xx + zeta_y * tau_xy + zeta_z * tau_xz;
        hv[p + 2 * sz] = zeta_x * tau_xy + zeta_y * tau_yy + zeta_z * tau_yz;
        hv[p + 3 * sz] = zeta_x * tau_xz + zeta_y * tau_yz + zeta_z * tau_zz;
        const double t_x = t_xi * xi_x + t_eta * eta_x + t_zeta * zeta_x;
        const double t_y = t_xi * xi_y + t_eta * eta_y + t_zeta * zeta_y;
        const double t_z = t_xi * xi_z + t_eta * eta_z + t_zeta * zeta_z;
        double conductivity;
        if (r->mixture) {
          conductivity = z->kappa[p];
        } else {
          const double cp = gamma_air * R_AIR / (gamma_air - 1);
          conductivity = mul / r->prandtl * cp;
        }
        const double u = pv[p + sz], v = pv[p + 2 * sz], w = pv[p + 3 * sz];
        const double Ex = u * tau_xx + v * tau_xy + w * tau_xz + conductivity * t_x;
        const double Ey = u * tau_xy + v * tau_yy + w * tau_yz + conductivity * t_y;
        const double Ez = u * tau_xz + v * tau_yz + w * tau_zz + conductivity * t_z;
        fv[p + 4 * sz] = xi_x * Ex + xi_y * Ey + xi_z * Ez;
        gv[p + 4 * sz] = eta_x * Ex + eta_y * Ey + eta_z * Ez;
        hv[p + 4 * sz] = zeta_x * Ex + zeta_y * Ey + zeta_z * Ez;
        /* two temperatures: kappa_ve grad Tve in the energy flux, src/ViscousScheme.cu:570-584;
         * the same conduction starts the E_ve flux, 703-719 */
        double eve_fx = 0, eve_fy = 0, eve_fz = 0;
        if (r->two_temperature) {
          const double tve_xi = DXI(z->tve, 0), tve_eta = DETA(z->tve, 0), tve_zeta = DZETA(z->tve, 0);
          const double tve_x = tve_xi * xi_x + tve_eta * eta_x + tve_zeta * zeta_x;
          const double tve_y = tve_xi * xi_y + tve_eta * eta_y + tve_zeta * zeta_y;
          const double tve_z = tve_xi * xi_z + tve_eta * eta_z + tve_zeta * zeta_z;
          const double kve = z->kappa_ve[p];
          eve_fx = kve * (xi_x * tve_x + xi_y * tve_y + xi_z * tve_z);
          eve_fy = kve * (eta_x * tve_x + eta_y * tve_y + eta_z * tve_z);
          eve_fz = kve * (zeta_x * tve_x + zeta_y * tve_y + zeta_z * tve_z);
          fv[p + 4 * sz] += eve_fx;
          gv[p + 4 * sz] += eve_fy;
          hv[p + 4 * sz] += eve_fz;
        }

        if (!r->mixture) continue;
        /* compute_viscous_flux_collocated_scalar<8>, src/ViscousScheme.cu:587-788 */
        const double *yv = z->sv;
        double diffusivity[OC_MAX_SPEC], yk[OC_MAX_SPEC];
        double dfx[OC_MAX_SPEC], dfy[OC_MAX_SPEC], dfz[OC_MAX_SPEC];
        double sGx = 0, sGy = 0, sGz = 0, sum_rhoDkYk = 0, mw_tot = 0;
        double CVT[3] = {0, 0, 0};
        for (int l = 0; l < ns; ++l) {
          yk[l] = yv[p + l * sz];
          diffusivity[l] = z->rhoD[p + l * sz];
          const double y_xi = DXI(yv, l), y_eta = DETA(yv, l), y_zeta = DZETA(yv, l);
          const double y_x = y_xi * xi_x + y_eta * eta_x + y_zeta * zeta_x;
          const double y_y = y_xi * xi_y + y_eta * eta_y + y_zeta * zeta_y;
          const double y_z = y_xi * xi_z + y_eta * eta_z + y_zeta * zeta_z;
          const double GX = xi_x * y_x + xi_y * y_y + xi_z * y_z;
          const double GE = eta_x * y_x + eta_y * y_y + eta_z * y_z;
          const double GZ = zeta_x * y_x + zeta_y * y_y + zeta_z * y_z;
          dfx[l] = GX;
          dfy[l] = GE;
          dfz[l] = GZ;
          CVT[0] += diffusivity[l] * GX;
          CVT[1] += diffusivity[l] * GE;
          CVT[2] += diffusivity[l] * GZ;
          sGx += GX * tb->imw[l];
          sGy += GE * tb->imw[l];
          sGz += GZ * tb->imw[l];
          mw_tot += yk[l] * tb->imw[l];
          sum_rhoDkYk += diffusivity[l] * yk[l];
        }
        mw_tot = 1.0 / mw_tot;
        CVT[0] -= mw_tot * sum_rhoDkYk * sGx;
        CVT[1] -= mw_tot * sum_rhoDkYk * sGy;
        CVT[2] -= mw_tot * sum_rhoDkYk * sGz;
        if (r->gradp_in_diffusion) {
          const double p_xi = DXI(pv, 4), p_eta = DETA(pv, 4), p_zeta = DZETA(pv, 4);
          const double p_x = p_xi * xi_x + p_eta * eta_x + p_zeta * zeta_x;
          const double p_y = p_xi * xi_y + p_eta * eta_y + p_zeta * zeta_y;
          const double p_z = p_xi * xi_z + p_eta * eta_z + p_zeta * zeta_z;
          const double pp = pv[p + 4 * sz];
          const double gX = (xi_x * p_x + xi_y * p_y + xi_z * p_z) / pp;
          const double gE = (eta_x * p_x + eta_y * p_y + eta_z * p_z) / pp;
          const double gZ = (zeta_x * p_x + zeta_y * p_y + zeta_z * p_z) / pp;
          for (int l = 0; l < ns; ++l) {
            const double coefficient = (mw_tot * tb->imw[l] - 1) * yk[l];
            dfx[l] += coefficient * gX;
            dfy[l] += coefficient * gE;
            dfz[l] += coefficient * gZ;
            CVT[0] += coefficient * gX * diffusivity[l];
            CVT[1] += coefficient * gE * diffusivity[l];
            CVT[2] += coefficient * gZ * diffusivity[l];
          }
        }
        double h[OC_MAX_SPEC];
        const double tm = pv[p + 5 * sz];
        oc_enthalpy(tb, tm, h);
        const double tve_m = r->two_temperature ? z->tve[p] : tm;
        for (int l = 0; l < ns; ++l) {
          /* compute_nonequilibrium_diffusion_enthalpy (src/Thermo.cuh:43-47) and the
           * E_ve carried by the diffusion flux, src/ViscousScheme.cu:722-756 */
          double h_diff = h[l], eve_l = 0;
          if (r->two_temperature) {
            eve_l = oc_ve_energy(tb, l, tve_m);
            h_diff = h[l] - oc_ve_energy(tb, l, tm) + eve_l;
          }
          double dflux = diffusivity[l] * (dfx[l] - mw_tot * yk[l] * sGx) - yk[l] * CVT[0];
          fv[p + (5 + l) * sz] = dflux;
          fv[p + 4 * sz] += h_diff * dflux;
          eve_fx += eve_l * dflux;
          dflux = diffusivity[l] * (dfy[l] - mw_tot * yk[l] * sGy) - yk[l] * CVT[1];
          gv[p + (5 + l) * sz] = dflux;
          gv[p + 4 * sz] += h_diff * dflux;
          eve_fy += eve_l * dflux;
          dflux = diffusivity[l] * (dfz[l] - mw_tot * yk[l] * sGz) - yk[l] * CVT[2];
          hv[p + (5 + l) * sz] = dflux;
          hv[p + 4 * sz] += h_diff * dflux;
          eve_fz += eve_l * dflux;
        }
        if (r->two_temperature) {
          fv[p + (5 + ns) * sz] = eve_fx;
          gv[p + (5 + ns) * sz] = eve_fy;
          hv[p + (5 + ns) * sz] = eve_fz;
        }
      }
}

/* compute_viscous_flux_derivative<8>, src/ViscousScheme.cu:790-834 */
static void viscous8_derivative(oc_zone *z, const oc_run *r) {
  const lay_t L = layout(z);
  const int mx = z->mx, my = z->my, mz = z->mz, nv = r->n_var;
  const long sz = L.sz, si = 1, sj = L.sx, sk = L.sxy;
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        int ct[6];
        btype_flags(z, i, j, k, ct);
        const long p = IX(L, i, j, k), q = DQX(z, i, j, k);
        const double jac = z->jac[p];
        for (int l = 1; l < nv; ++l)
          z->dq[q + l * dsz] += (DXI(z->ff, l) + DETA(z->gf, l) + DZETA(z->hf, l)) * jac;
      }
#undef DXI
#undef DETA
#undef DZETA
}

/* Index of the target node for loop position n of an inner face, and the
 * source node (src/DataCommunication.cu:43-52). */
static void inner_map(const oc_inner_face *f, const int n[3], int idx[3], int idx_tar[3]) {
  int d_idx[3];
  for (int a = 0; a < 3; ++a) {
    d_idx[a] = f->loop_dir[a] * n[a];
    idx[a] = f->range_start[a] + d_idx[a];
  }
  for (int a = 0; a < 3; ++a) idx_tar[a] = f->target_start[a] + f->target_loop_dir[a] * d_idx[f->src_tar[a]];
}

/* inner_exchange(task 1), src/DataCommunication.cu:489-541: coincident nodes
 * average f/g/hFlux between the blocks, ghosts take the neighbour's values. */
static void inner_exchange_flux(oc_zone *z, oc_zone *tz, const oc_inner_face *f, const oc_run *r) {
  const lay_t L = layout(z), T = layout(tz);
  double *a[3] = {z->ff, z->gf, z->hf}, *b[3] = {tz->ff, tz->gf, tz->hf};
  const int face_dir = f->direction > 0 ? f->range_start[f->face] : f->range_end[f->face];
  for (int n2 = 0; n2 < f->n_point[2]; ++n2)
    for (int n1 = 0; n1 < f->n_point[1]; ++n1)
      for (int n0 = 0; n0 < f->n_point[0]; ++n0) {
        const int n[3] = {n0, n1, n2};
        int idx[3], it[3];
        inner_map(f, n, idx, it);
        const long p = IX(L, idx[0], idx[1], idx[2]), q = IX(T, it[0], it[1], it[2]);
        for (int l = 0; l < r->n_var; ++l)
          for (int c = 0; c < 3; ++c) {
            if (idx[f->face] == face_dir) {
              const double ave = 0.5 * (a[c][p + l * L.sz] + b[c][q + l * T.sz]);
              a[c][p + l * L.sz] = ave;
              b[c][q + l * T.sz] = ave;
            } else {
              a[c][p + l * L.sz] = b[c][q + l * T.sz];
            }
          }
      }
}

/* compute_{f,g,h}v_2nd_order + compute_d{F,G,H}v_d{x,y,z} (one temperature),
 * src/ViscousScheme.cu:86-138, 836-1512: the flux on the face between node P
 * and P+e_dir from first differences across the face and four-point central
 * differences along the other axes, face-averaged mu, kappa, rho*D, metrics
 * and J*grad(dir); then dq[1..] += F(face) - F(face - 1).  fFlux is the
 * scratch plane set for the faces (components 0..nv-2). */
static void viscous2_dir(oc_zone *z, const oc_run *r, const oc_tables *tb, int dir) {
  const lay_t L = layout(z);
  const long sz = L.sz;
  const int nv = r->n_var, ns = r->n_spec;
  const int n[3] = {z->mx, z->my, z->mz};
  const long st[3] = {1, L.sx, L.sxy};
  double *fv = z->ff;
  const double *pv = z->bv, *mt = z->metric;
#pragma omp parallel for schedule(static)
  for (int k = (dir == 2 ? -1 : 0); k < n[2]; ++k)
    for (int j = (dir == 1 ? -1 : 0); j < n[1]; ++j)
      for (int i = (dir == 0 ? -1 : 0); i < n[0]; ++i) {
        const long p = IX(L, i, j, k), pe = p + st[dir];
        double m[9];
        for (int a = 0; a < 9; ++a) m[a] = 0.5 * (mt[p + a * sz] + mt[pe + a * sz]);
        /* d/dxi, d/deta, d/dzeta of plane f at the face */
#define DD(arr, l, a)                                                                                   \
  ((a) == dir ? (arr)[pe + (long)(l) * sz] - (arr)[p + (long)(l) * sz]                                 \
              : 0.25 * ((arr)[p + st[a] + (long)(l) * sz] - (arr)[p - st[a] + (long)(l) * sz] +        \
                        (arr)[pe + st[a] + (long)(l) * sz] - (arr)[pe - st[a] + (long)(l) * sz]))
        double g[4][3]; /* u, v, w, T gradients in x, y, z */
        const int var[4] = {1, 2, 3, 5};
        for (int q = 0; q < 4; ++q) {
          const double d0 = DD(pv, var[q], 0), d1 = DD(pv, var[q], 1), d2 = DD(pv, var[q], 2);
          g[q][0] = d0 * m[0] + d1 * m[3] + d2 * m[6];
          g[q][1] = d0 * m[1] + d1 * m[4] + d2 * m[7];
          g[q][2] = d0 * m[2] + d1 * m[5] + d2 * m[8];
        }
        const double mul = 0.5 * (z->mul[p] + z->mul[pe]);
        const double u_x = g[0][0], u_y = g[0][1], u_z = g[0][2], v_x = g[1][0], v_y = g[1][1], v_z = g[1][2];
        const double w_x = g[2][0], w_y = g[2][1], w_z = g[2][2];
        const double tau_xx = mul * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
        const double tau_yy = mul * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
        const double tau_zz = mul * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
        const double tau_xy = mul * (u_y + v_x), tau_xz = mul * (u_z + w_x), tau_yz = mul * (v_z + w_y);
        const double kx = 0.5 * (mt[p + (3 * dir) * sz] * z->jac[p] + mt[pe + (3 * dir) * sz] * z->jac[pe]);
        const double ky = 0.5 * (mt[p + (3 * dir + 1) * sz] * z->jac[p] + mt[pe + (3 * dir + 1) * sz] * z->jac[pe]);
        const double kz = 0.5 * (mt[p + (3 * dir + 2) * sz] * z->jac[p] + mt[pe + (3 * dir + 2) * sz] * z->jac[pe]);
        fv[p] = kx * tau_xx + ky * tau_xy + kz * tau_xz;
        fv[p + sz] = kx * tau_xy + ky * tau_yy + kz * tau_yz;
        fv[p + 2 * sz] = kx * tau_xz + ky * tau_yz + kz * tau_zz;
        const double um = 0.5 * (pv[p + sz] + pv[pe + sz]), vm = 0.5 * (pv[p + 2 * sz] + pv[pe + 2 * sz]),
                     wm = 0.5 * (pv[p + 3 * sz] + pv[pe + 3 * sz]);
        const double conductivity = r->mixture ? 0.5 * (z->kappa[p] + z->kappa[pe])
                                               : mul / r->prandtl * (gamma_air * R_AIR / (gamma_air - 1));
        fv[p + 3 * sz] = um * fv[p] + vm * fv[p + sz] + wm * fv[p + 2 * sz] +
                         conductivity * (kx * g[3][0] + ky * g[3][1] + kz * g[3][2]);
        if (r->mixture) {
          const double *y = z->sv;
          double diff[OC_MAX_SPEC], yk[OC_MAX_SPEC], force[OC_MAX_SPEC];
          double sGw = 0, sDY = 0, CV = 0, mw_tot = 0;
          for (int l = 0; l < ns; ++l) {
            yk[l] = 0.5 * (y[p + l * sz] + y[pe + l * sz]);
            diff[l] = 0.5 * (z->rhoD[p + l * sz] + z->rhoD[pe + l * sz]);
            const double d0 = DD(y, l, 0), d1 = DD(y, l, 1), d2 = DD(y, l, 2);
            const double y_x = d0 * m[0] + d1 * m[3] + d2 * m[6];
            const double y_y = d0 * m[1] + d1 * m[4] + d2 * m[7];
            const double y_z = d0 * m[2] + d1 * m[5] + d2 * m[8];
            const double G = kx * y_x + ky * y_y + kz * y_z;
            force[l] = G;
            CV += diff[l] * G;
            sGw += G * tb->imw[l];
            mw_tot += yk[l] * tb->imw[l];
            sDY += diff[l] * yk[l];
          }
          mw_tot = 1.0 / mw_tot;
          CV -= mw_tot * sDY * sGw;
          if (r->gradp_in_diffusion) {
            const double d0 = DD(pv, 4, 0), d1 = DD(pv, 4, 1), d2 = DD(pv, 4, 2);
            const double p_x = d0 * m[0] + d1 * m[3] + d2 * m[6];
            const double p_y = d0 * m[1] + d1 * m[4] + d2 * m[7];
            const double p_z = d0 * m[2] + d1 * m[5] + d2 * m[8];
            const double gp = (kx * p_x + ky * p_y + kz * p_z) / (0.5 * (pv[pe + 4 * sz] + pv[p + 4 * sz]));
            for (int l = 0; l < ns; ++l) {
              force[l] += (mw_tot * tb->imw[l] - 1) * yk[l] * gp;
              CV += (mw_tot * tb->imw[l] - 1) * yk[l] * gp * diff[l];
            }
          }
          double h[OC_MAX_SPEC];
          const double tm = 0.5 * (pv[p + 5 * sz] + pv[pe + 5 * sz]);
          oc_enthalpy(tb, tm, h);
          const double tve_m = r->two_temperature ? 0.5 * (z->tve[p] + z->tve[pe]) : tm;
          for (int l = 0; l < ns; ++l) {
            const double dflux = diff[l] * (force[l] - mw_tot * yk[l] * sGw) - yk[l] * CV;
            fv[p + (4 + l) * sz] = dflux;
            /* compute_nonequilibrium_diffusion_enthalpy, src/Thermo.cuh:43-47 */
            const double h_diff =
                r->two_temperature ? h[l] - oc_ve_energy(tb, l, tm) + oc_ve_energy(tb, l, tve_m) : h[l];
            fv[p + 3 * sz] += h_diff * dflux;
          }
          if (r->two_temperature) {
            /* kappa_ve grad Tve conduction and the E_ve flux, src/ViscousScheme.cu:1040-1055 */
            const double d0 = DD(z->tve, 0, 0), d1 = DD(z->tve, 0, 1), d2 = DD(z->tve, 0, 2);
            const double tve_x = d0 * m[0] + d1 * m[3] + d2 * m[6];
            const double tve_y = d0 * m[1] + d1 * m[4] + d2 * m[7];
            const double tve_z = d0 * m[2] + d1 * m[5] + d2 * m[8];
            const double k_ve = 0.5 * (z->kappa_ve[p] + z->kappa_ve[pe]);
            const double cond = k_ve * (kx * tve_x + ky * tve_y + kz * tve_z);
            fv[p + 3 * sz] += cond;
            double eve_flux = cond;
            for (int l = 0; l < ns; ++l) eve_flux += oc_ve_energy(tb, l, tve_m) * fv[p + (4 + l) * sz];
            fv[p + (4 + ns) * sz] = eve_flux;
          }
        }
#undef DD
      }
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k), q = DQX(z, i, j, k);
        for (int l = 1; l < nv; ++l) z->dq[q + l * dsz] += fv[p + (l - 1) * sz] - fv[p - st[dir] + (l - 1) * sz];
      }
}

/* compute_viscous_flux, 2nd-order branch, src/ViscousScheme.cu:32-60 */
void oc_viscous2(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  viscous2_dir(z, r, tb, 0);
  viscous2_dir(z, r, tb, 1);
  if (z->mz > 1) viscous2_dir(z, r, tb, 2);
}

void oc_viscous8(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  oc_zone *zs[1] = {z};
  oc_viscous8_blocks(zs, 1, r, tb);
}

/* compute_viscous_flux, 8th-order branch, src/ViscousScheme.cu:62-83: nodal
 * fluxes on every block, exchange_value(task 1) (inner faces of every block,
 * then periodic faces of every block, src/DataCommunication.cu:437-485), then
 * the derivative on every block. */
void oc_viscous8_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb) {
  for (int b = 0; b < n_block; ++b) viscous8_flux(zones[b], r, tb);
  for (int b = 0; b < n_block; ++b)
    for (int f = 0; f < zones[b]->n_inner; ++f) {
      const oc_inner_face *fc = &zones[b]->inner[f];
      inner_exchange_flux(zones[b], zones[fc->target_block], fc, r);
    }
  for (int b = 0; b < n_block; ++b)
    for (int f = 0; f < zones[b]->n_boundary; ++f)
      if (zones[b]->boundary[f].type == 8) periodic_exchange_flux(zones[b], r, &zones[b]->boundary[f]);
  for (int b = 0; b < n_block; ++b) viscous8_derivative(zones[b], r);
}

static int sgn_int(int a) { return a < 0 ? -1 : 1; } /* gxl::sgn, src/gxl_lib/Math.hpp:10-12 */

/* InnerFace::register_boundary, src/Mesh.cpp:39-172.  The file's numbers are
 * 1-based; the face-normal axis repeats one value (1: the low face), one
 * tangential axis of each side is negated to mark the axes that correspond,
 * and the order of start/end gives the running direction. */
void oc_register_inner(const int src[6], const int tar[6], int target_block, int dim, int ngg, oc_inner_face *f) {
  int rs[3] = {src[0], src[2], src[4]}, re[3] = {src[1], src[3], src[5]};
  int ts[3] = {tar[0], tar[2], tar[4]}, te[3] = {tar[1], tar[3], tar[5]};
  int st[3] = {0, 0, 0};
  int face = 0, dirn = 1, tface = 0, tdirn = 1;
  for (int a = 0; a < dim; ++a)
    if (rs[a] == re[a]) {
      face = a;
      if (rs[a] == 1) dirn = -1;
      --rs[a];
      --re[a];
      break;
    }
  for (int a = 0; a < dim; ++a)
    if (ts[a] == te[a]) {
      tface = a;
      if (ts[a] == 1) tdirn = -1;
      --ts[a];
      --te[a];
      break;
    }
  st[tface] = -(dirn * tdirn) * (face + 1);
  if (dim == 2) {
    const int nf = 1 - face, tnf = 1 - tface;
    rs[nf] = abs(rs[nf]) - 1;
    re[nf] = abs(re[nf]) - 1;
    ts[tnf] = abs(ts[tnf]) - 1;
    te[tnf] = abs(te[tnf]) - 1;
    st[tnf] = sgn_int((te[tnf] - ts[tnf]) * (re[nf] - rs[nf])) * (nf + 1);
    st[2] = 3;
    rs[2] = re[2] = ts[2] = te[2] = 0;
  } else {
    int nf = 0, tnf = 0;
    for (int a = 0; a < dim; ++a)
      if (rs[a] < 0) {
        nf = a;
        rs[a] = -rs[a] - 1;
        re[a] = -re[a] - 1;
        break;
      }
    for (int a = 0; a < dim; ++a)
      if (ts[a] < 0) {
        tnf = a;
        ts[a] = -ts[a] - 1;
        te[a] = -te[a] - 1;
        break;
      }
    st[tnf] = sgn_int((te[tnf] - ts[tnf]) * (re[nf] - rs[nf])) * (nf + 1);
    nf = 3 - face - nf;
    tnf = 3 - tface - tnf;
    --rs[nf];
    --re[nf];
    --ts[tnf];
    --te[tnf];
    st[tnf] = sgn_int((te[tnf] - ts[tnf]) * (re[nf] - rs[nf])) * (nf + 1);
  }
  /* the loop along the normal covers the face node and the ghost layers */
  rs[face] = dirn < 0 ? -ngg : re[face];
  re[face] = dirn < 0 ? 0 : re[face] + ngg;
  const int sgn_tar = sgn_int(st[tface]);
  if (tdirn > 0)
    ts[tface] = sgn_tar > 0 ? te[tface] - ngg : te[tface];
  else
    ts[tface] = sgn_tar > 0 ? 0 : ngg;
  for (int a = 0; a < 3; ++a) {
    f->range_start[a] = rs[a];
    f->range_end[a] = re[a];
    f->target_start[a] = ts[a];
    f->loop_dir[a] = rs[a] <= re[a] ? 1 : -1;
    f->target_loop_dir[a] = sgn_int(st[a]);
    f->src_tar[a] = abs(st[a]) - 1;
    f->n_point[a] = abs(rs[a] - re[a]) + 1;
  }
  f->face = face;
  f->direction = dirn;
  f->target_face = tface;
  f->target_direction = tdirn;
  f->target_block = target_block - 1;
}

/* inner_communication, src/DataCommunication.cu:37-124.  Coincident nodes
 * take the average of both blocks (or the neighbour's values where this side
 * moves and the neighbour does not), ghost nodes the neighbour's interior;
 * cv is recomputed on every node written. */
static void inner_communication(oc_zone *z, oc_zone *tz, const oc_inner_face *f, const oc_run *r,
                                const oc_tables *tb) {
  const lay_t L = layout(z), T = layout(tz);
  const int face_dir = f->direction > 0 ? f->range_start[f->face] : f->range_end[f->face];
  for (int n2 = 0; n2 < f->n_point[2]; ++n2)
    for (int n1 = 0; n1 < f->n_point[1]; ++n1)
      for (int n0 = 0; n0 < f->n_point[0]; ++n0) {
        const int n[3] = {n0, n1, n2};
        int idx[3], it[3];
        inner_map(f, n, idx, it);
        const long p = IX(L, idx[0], idx[1], idx[2]), q = IX(T, it[0], it[1], it[2]);
        if (idx[f->face] == face_dir) {
          const double u_this = norm3(z->bv[p + L.sz], z->bv[p + 2 * L.sz], z->bv[p + 3 * L.sz]);
          if (fabs(u_this) > 1e-15) {
            const double u_that = norm3(tz->bv[q + T.sz], tz->bv[q + 2 * T.sz], tz->bv[q + 3 * T.sz]);
            if (fabs(u_that) < 1e-15) {
              for (int l = 0; l < 6; ++l) z->bv[p + l * L.sz] = tz->bv[q + l * T.sz];
              for (int l = 0; l < r->n_scalar; ++l) z->sv[p + l * L.sz] = tz->sv[q + l * T.sz];
              if (r->two_temperature) z->tve[p] = tz->tve[q];
            } else {
              for (int l = 0; l < 6; ++l) {
                const double ave = 0.5 * (tz->bv[q + l * T.sz] + z->bv[p + l * L.sz]);
                z->bv[p + l * L.sz] = ave;
                tz->bv[q + l * T.sz] = ave;
              }
              for (int l = 0; l < r->n_scalar; ++l) {
                const double ave = 0.5 * (tz->sv[q + l * T.sz] + z->sv[p + l * L.sz]);
                z->sv[p + l * L.sz] = ave;
                tz->sv[q + l * T.sz] = ave;
              }
              if (r->two_temperature) {
                const double ave = 0.5 * (tz->tve[q] + z->tve[p]);
                z->tve[p] = ave;
                tz->tve[q] = ave;
              }
              cv_from_bv_point(tz, &T, r, tb, q);
            }
          }
        } else {
          for (int l = 0; l < 6; ++l) z->bv[p + l * L.sz] = tz->bv[q + l * T.sz];
          for (int l = 0; l < r->n_scalar; ++l) z->sv[p + l * L.sz] = tz->sv[q + l * T.sz];
          if (r->two_temperature) z->tve[p] = tz->tve[q];
        }
        cv_from_bv_point(z, &L, r, tb, p);
      }
}

void oc_data_communication(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb) {
  for (int b = 0; b < n_block; ++b)
    for (int f = 0; f < zones[b]->n_inner; ++f) {
      const oc_inner_face *fc = &zones[b]->inner[f];
      inner_communication(zones[b], zones[fc->target_block], fc, r, tb);
    }
}

/* ------------------------------------------------------------------------ */
/* boundary conditions                                                       */
/* ------------------------------------------------------------------------ */
/* apply_outflow (plain zero-gradient branch), src/BoundCond.cu:2204-2349 */
static void bc_outflow(oc_zone *z, const oc_run *r, const oc_tables *tb, const oc_boundary *b) {
  const lay_t L = layout(z);
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        const long p = IX(L, i, j, k);
        for (int g = 1; g <= z->ngg; ++g) {
          const long pg = IX(L, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          for (int l = 0; l < 6; ++l) z->bv[pg + l * L.sz] = z->bv[p + l * L.sz];
          for (int l = 0; l < r->n_scalar; ++l) z->sv[pg + l * L.sz] = z->sv[p + l * L.sz];
          if (r->two_temperature) z->tve[pg] = z->tve[p] > 0.0 ? z->tve[p] : z->bv[p + 5 * L.sz];
          cv_from_bv_point(z, &L, r, tb, pg);
        }
      }
}

/* apply_periodic, src/BoundCond.cu:3389-3440 */
static void bc_periodic(oc_zone *z, const oc_run *r, const oc_tables *tb, const oc_boundary *b) {
  const lay_t L = layout(z);
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  const int nn[3] = {z->mx, z->my, z->mz};
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        int o[3] = {i, j, k};
        o[b->face] = b->direction < 0 ? nn[b->face] - 1 : 0;
        const long p = IX(L, i, j, k), po = IX(L, o[0], o[1], o[2]);
        for (int l = 0; l < 6; ++l) {
          const double ave = 0.5 * (z->bv[p + l * L.sz] + z->bv[po + l * L.sz]);
          z->bv[p + l * L.sz] = ave;
          z->bv[po + l * L.sz] = ave;
        }
        for (int l = 0; l < r->n_scalar; ++l) {
          const double ave = 0.5 * (z->sv[p + l * L.sz] + z->sv[po + l * L.sz]);
          z->sv[p + l * L.sz] = ave;
          z->sv[po + l * L.sz] = ave;
        }
        cv_from_bv_point(z, &L, r, tb, p);
        for (int g = 1; g <= z->ngg; ++g) {
          const long pg = IX(L, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          const long pi = IX(L, o[0] + g * dir[0], o[1] + g * dir[1], o[2] + g * dir[2]);
          for (int l = 0; l < 6; ++l) z->bv[pg + l * L.sz] = z->bv[pi + l * L.sz];
          for (int l = 0; l < r->n_scalar; ++l) z->sv[pg + l * L.sz] = z->sv[pi + l * L.sz];
          cv_from_bv_point(z, &L, r, tb, pg);
        }
      }
}

static double mixture_r(const oc_run *r, const oc_tables *tb, const double *sv, long p, long sz) {
  if (!r->mixture) return R_AIR;
  double R = 0;
  for (int l = 0; l < r->n_spec; ++l) R += sv[p + l * sz] * tb->gas_const[l];
  return R;
}

/* apply_symmetry, src/BoundCond.cu:2130-2202: the face node keeps the
 * tangential velocity of its interior neighbour and its rho, p, T, scalars;
 * ghosts mirror the interior with the normal velocity reversed. */
static void bc_symmetry(oc_zone *z, const oc_run *r, const oc_tables *tb, const oc_boundary *b) {
  const lay_t L = layout(z);
  const long sz = L.sz;
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        const long p = IX(L, i, j, k), pin = IX(L, i - dir[0], j - dir[1], k - dir[2]);
        double kx = z->metric[p + (3 * b->face) * sz], ky = z->metric[p + (3 * b->face + 1) * sz],
               kz = z->metric[p + (3 * b->face + 2) * sz];
        const double km = sqrt(kx * kx + ky * ky + kz * kz);
        kx /= km;
        ky /= km;
        kz /= km;
        const double u1 = z->bv[pin + sz], v1 = z->bv[pin + 2 * sz], w1 = z->bv[pin + 3 * sz];
        double u_k = kx * u1 + ky * v1 + kz * w1;
        z->bv[p + sz] = u1 - kx * u_k;
        z->bv[p + 2 * sz] = v1 - ky * u_k;
        z->bv[p + 3 * sz] = w1 - kz * u_k;
        z->bv[p] = z->bv[pin];
        z->bv[p + 4 * sz] = z->bv[pin + 4 * sz];
        z->bv[p + 5 * sz] = z->bv[pin + 5 * sz];
        for (int l = 0; l < r->n_scalar; ++l) z->sv[p + l * sz] = z->sv[pin + l * sz];
        if (r->two_temperature) z->tve[p] = z->tve[pin];
        cv_from_bv_point(z, &L, r, tb, p);
        for (int g = 1; g <= z->ngg; ++g) {
          const long pg = IX(L, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          const long pi = IX(L, i - g * dir[0], j - g * dir[1], k - g * dir[2]);
          z->bv[pg] = z->bv[pi];
          const double u = z->bv[pi + sz], v = z->bv[pi + 2 * sz], w = z->bv[pi + 3 * sz];
          u_k = kx * u + ky * v + kz * w;
          z->bv[pg + sz] = u - 2 * u_k * kx;
          z->bv[pg + 2 * sz] = v - 2 * u_k * ky;
          z->bv[pg + 3 * sz] = w - 2 * u_k * kz;
          z->bv[pg + 4 * sz] = z->bv[pi + 4 * sz];
          z->bv[pg + 5 * sz] = z->bv[pi + 5 * sz];
          for (int l = 0; l < r->n_scalar; ++l) z->sv[pg + l * sz] = z->sv[pi + l * sz];
          if (r->two_temperature) z->tve[pg] = z->tve[pi];
          cv_from_bv_point(z, &L, r, tb, pg);
        }
      }
}

/* apply_inflow, constant inflow (inflow_type 0, no fluctuation),
 * src/BoundCond.cu:2810-2997: the face node takes the inflow state, or with
 * inflow_sub the characteristic combination of it with the interior
 * neighbour unless the flow at the face is supersonic inward (Mn < -1);
 * ghosts take the face state, or with inflow_sub its linear extrapolation. */
static void bc_inflow(oc_zone *z, const oc_run *r, const oc_tables *tb, const oc_boundary *b) {
  const lay_t L = layout(z);
  const long sz = L.sz;
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  /* the inflow scalars; with two temperatures E_ve of the stream at its Tve (Inflow's constructor,
   * src/BoundCond.cpp:256-262) */
  double svb[OC_MAX_SPEC];
  for (int l = 0; l < r->n_scalar; ++l) svb[l] = b->sv[l];
  if (r->two_temperature) {
    const double tve = b->temperature_ve > 0 ? b->temperature_ve : b->temperature;
    double eve = 0;
    for (int l = 0; l < r->n_spec; ++l) eve += b->sv[l] * oc_ve_energy(tb, l, tve);
    svb[r->n_spec] = eve;
  }
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        const long p = IX(L, i, j, k);
        double density = b->density, u = b->u, v = b->v, w = b->w, pr = b->pressure, T = b->temperature;
        if (b->inflow_sub) {
          double nx = z->metric[p + (3 * b->face) * sz], ny = z->metric[p + (3 * b->face + 1) * sz],
                 nz = z->metric[p + (3 * b->face + 2) * sz];
          const double gn = b->direction / sqrt(nx * nx + ny * ny + nz * nz);
          nx *= gn;
          ny *= gn;
          nz *= gn;
          const double c0 = r->mixture ? z->acoustic[p] : sqrt(gamma_air * R_AIR * z->bv[p + 5 * sz]);
          const double Mn = (z->bv[p + sz] * nx + z->bv[p + 2 * sz] * ny + z->bv[p + 3 * sz] * nz) / c0;
          if (!(Mn < -1.0)) {
            const long p1 = IX(L, i - dir[0], j - dir[1], k - dir[2]);
            const double c1 = r->mixture ? z->acoustic[p1] : sqrt(gamma_air * R_AIR * z->bv[p1 + 5 * sz]);
            const double rho1 = z->bv[p1], p1v = z->bv[p1 + 4 * sz];
            pr = 0.5 * (p1v + b->pressure -
                        rho1 * c1 * ((b->u - z->bv[p1 + sz]) * nx + (b->v - z->bv[p1 + 2 * sz]) * ny +
                                     (b->w - z->bv[p1 + 3 * sz]) * nz));
            density = b->density + (pr - b->pressure) / (c1 * c1);
            u = b->u - (pr - b->pressure) / (rho1 * c1) * nx;
            v = b->v - (pr - b->pressure) / (rho1 * c1) * ny;
            w = b->w - (pr - b->pressure) / (rho1 * c1) * nz;
            T = r->mixture ? pr * b->mw / (density * R_u) : pr / (density * R_AIR);
          }
        }
        z->bv[p] = density;
        z->bv[p + sz] = u;
        z->bv[p + 2 * sz] = v;
        z->bv[p + 3 * sz] = w;
        z->bv[p + 4 * sz] = pr;
        z->bv[p + 5 * sz] = T;
        for (int l = 0; l < r->n_scalar; ++l) z->sv[p + l * sz] = svb[l];
        cv_from_bv_point(z, &L, r, tb, p);
        for (int g = 1; g <= z->ngg; ++g) {
          const long pg = IX(L, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          for (int l = 0; l < r->n_scalar; ++l) z->sv[pg + l * sz] = svb[l];
          if (b->inflow_sub) {
            const long pi = IX(L, i - g * dir[0], j - g * dir[1], k - g * dir[2]);
            double rho2 = 2 * density - z->bv[pi], p2 = 2 * pr - z->bv[pi + 4 * sz];
            if (p2 <= 0.1 * pr) p2 = 0.1 * pr;
            if (rho2 <= 0.1 * density) rho2 = 0.1 * density;
            z->bv[pg] = rho2;
            z->bv[pg + sz] = 2 * u - z->bv[pi + sz];
            z->bv[pg + 2 * sz] = 2 * v - z->bv[pi + 2 * sz];
            z->bv[pg + 3 * sz] = 2 * w - z->bv[pi + 3 * sz];
            z->bv[pg + 4 * sz] = p2;
            z->bv[pg + 5 * sz] = r->mixture ? p2 * b->mw / (rho2 * R_u) : p2 / (rho2 * R_AIR);
          } else {
            z->bv[pg] = density;
            z->bv[pg + sz] = u;
            z->bv[pg + 2 * sz] = v;
            z->bv[pg + 3 * sz] = w;
            z->bv[pg + 4 * sz] = pr;
            z->bv[pg + 5 * sz] = T;
          }
          cv_from_bv_point(z, &L, r, tb, pg);
        }
      }
}

/* project_to_interior in apply_wall, src/BoundCond.cu:3070-3087: a reference
 * node on another physical face moves one node inward along that axis. */
static void project_to_interior(const oc_zone *z, int *ix) {
  if (ix[0] <= 0 && z->btype_il[ix[1] + ix[2] * z->my] != 0)
    ix[0] = 1;
  else if (ix[0] >= z->mx - 1 && z->btype_ir[ix[1] + ix[2] * z->my] != 0)
    ix[0] = z->mx - 2;
  if (ix[1] <= 0 && z->btype_jl[ix[0] + ix[2] * z->mx] != 0)
    ix[1] = 1;
  else if (ix[1] >= z->my - 1 && z->btype_jr[ix[0] + ix[2] * z->mx] != 0)
    ix[1] = z->my - 2;
  if (z->mz > 1) {
    if (ix[2] <= 0 && z->btype_kl[ix[0] + ix[1] * z->mx] != 0)
      ix[2] = 1;
    else if (ix[2] >= z->mz - 1 && z->btype_kr[ix[0] + ix[1] * z->mx] != 0)
      ix[2] = z->mz - 2;
  }
}

/* apply_wall (isothermal or adiabatic, non-catalytic, at rest),
 * src/BoundCond.cu:3055-3387: no-slip face node with the wall temperature
 * and the interior pressure; ghosts reflect the velocity and take
 * T_g = 2 T_w - T_i (isothermal, floored at T_w when below 0.1 T_w) or T_i. */
static void bc_wall(oc_zone *z, const oc_run *r, const oc_tables *tb, const oc_boundary *b) {
  const lay_t L = layout(z);
  const long sz = L.sz;
  int dir[3] = {0, 0, 0};
  dir[b->face] = b->direction;
  for (int k = b->range_start[2]; k <= b->range_end[2]; ++k)
    for (int j = b->range_start[1]; j <= b->range_end[1]; ++j)
      for (int i = b->range_start[0]; i <= b->range_end[0]; ++i) {
        const long p = IX(L, i, j, k);
        int ix[3] = {i - dir[0], j - dir[1], k - dir[2]};
        project_to_interior(z, ix);
        const long pin = IX(L, ix[0], ix[1], ix[2]);
        double t_wall = z->bv[p + 5 * sz];
        if (b->thermal_type == 0)
          t_wall = b->wall_temperature;
        else
          t_wall = z->bv[pin + 5 * sz];
        /* Tve of the wall: the interior's (adiabatic) or the condition's, else T_w (src/BoundCond.cu:3167-3177) */
        double tve_wall = t_wall;
        if (r->two_temperature) {
          const double tve_in = z->tve[pin] > 0.0 ? z->tve[pin] : z->bv[pin + 5 * sz];
          tve_wall = b->thermal_type == 1 ? tve_in : (b->temperature_ve > 0.0 ? b->temperature_ve : t_wall);
        }
        const double pr = z->bv[pin + 4 * sz];
        for (int l = 0; l < r->n_spec; ++l) z->sv[p + l * sz] = z->sv[pin + l * sz];
        const double R = mixture_r(r, tb, z->sv, p, sz);
        z->bv[p] = pr / (t_wall * R);
        z->bv[p + sz] = 0;
        z->bv[p + 2 * sz] = 0;
        z->bv[p + 3 * sz] = 0;
        z->bv[p + 4 * sz] = pr;
        z->bv[p + 5 * sz] = t_wall;
        if (r->two_temperature) {
          double y[OC_MAX_SPEC];
          for (int l = 0; l < r->n_spec; ++l) y[l] = z->sv[p + l * sz];
          z->tve[p] = tve_wall;
          z->sv[p + r->n_spec * sz] = oc_mixture_ve_energy(tb, tve_wall, y, NULL);
        }
        cv_from_bv_point(z, &L, r, tb, p);
        for (int g = 1; g <= z->ngg; ++g) {
          int in[3] = {i - g * dir[0], j - g * dir[1], k - g * dir[2]};
          project_to_interior(z, in);
          const long pi = IX(L, in[0], in[1], in[2]);
          const long pg = IX(L, i + g * dir[0], j + g * dir[1], k + g * dir[2]);
          const double p_i = z->bv[pi + 4 * sz], t_i = z->bv[pi + 5 * sz];
          double t_g = t_i;
          if (b->thermal_type == 0) {
            t_g = 2 * t_wall - t_i;
            /* floor 0.1 T_w in the two-temperature build, T_w otherwise (src/BoundCond.cu:3311-3318) */
            if (t_g <= 0.1 * t_wall) t_g = r->two_temperature ? 0.1 * t_wall : t_wall;
          }
          for (int l = 0; l < r->n_spec; ++l) z->sv[pg + l * sz] = z->sv[pi + l * sz];
          if (r->two_temperature) {
            const double tve_i = z->tve[pi] > 0.0 ? z->tve[pi] : t_i;
            double tve_g = tve_i;
            if (b->thermal_type == 0) {
              tve_g = 2 * tve_wall - tve_i;
              const double floor_ve = 0.1 * dmax(tve_wall, 1.0);
              if (tve_g <= floor_ve) tve_g = floor_ve;
            }
            double y[OC_MAX_SPEC];
            for (int l = 0; l < r->n_spec; ++l) y[l] = z->sv[pg + l * sz];
            z->tve[pg] = tve_g;
            z->sv[pg + r->n_spec * sz] = oc_mixture_ve_energy(tb, tve_g, y, NULL);
          }
          const double Rg = mixture_r(r, tb, z->sv, pi, sz);
          z->bv[pg] = p_i / (t_g * Rg);
          z->bv[pg + sz] = -z->bv[pi + sz];
          z->bv[pg + 2 * sz] = -z->bv[pi + 2 * sz];
          z->bv[pg + 3 * sz] = -z->bv[pi + 3 * sz];
          z->bv[pg + 4 * sz] = p_i;
          z->bv[pg + 5 * sz] = t_g;
          cv_from_bv_point(z, &L, r, tb, pg);
        }
      }
}

/* apply_boundary_conditions, src/BoundCond.cu:3846-3997, in its priority
 * order: outflow (6), inflow (5), symmetry (3), wall (2), periodic (8). */
void oc_boundary_conditions(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  for (int f = 0; f < z->n_boundary; ++f)
    if (z->boundary[f].type == 6) bc_outflow(z, r, tb, &z->boundary[f]);
  for (int f = 0; f < z->n_boundary; ++f)
    if (z->boundary[f].type == 5) bc_inflow(z, r, tb, &z->boundary[f]);
  for (int f = 0; f < z->n_boundary; ++f)
    if (z->boundary[f].type == 3) bc_symmetry(z, r, tb, &z->boundary[f]);
  for (int f = 0; f < z->n_boundary; ++f)
    if (z->boundary[f].type == 2) bc_wall(z, r, tb, &z->boundary[f]);
  for (int f = 0; f < z->n_boundary; ++f)
    if (z->boundary[f].type == 8) bc_periodic(z, r, tb, &z->boundary[f]);
}

/* limit_flow bounds, src/DParameter.cu:371-387; v_inf is the reference speed
 * already resolved (the sound speed where |v_inf| < 1). */
static void limit_bounds(const oc_run *r, double *ll, double *ul) {
  ll[0] = 1e-6 * r->rho_inf;
  ul[0] = 1e+3 * r->rho_inf;
  const double v = r->v_inf;
  for (int l = 1; l < 4; ++l) {
    ll[l] = -1000 * v;
    ul[l] = 1000 * v;
  }
  ll[4] = 1e-6 * r->p_inf;
  ul[4] = 1e+3 * r->p_inf;
}

static int point_bad(const double *bv, long p, long sz, const double *ll, const double *ul) {
  for (int l = 0; l < 5; ++l) {
    const double v = bv[p + l * sz];
    if (isnan(v) || v < ll[l] || v > ul[l]) return 1;
  }
  return 0;
}

void oc_limit_flow(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  const lay_t L = layout(z);
  const long sz = L.sz;
  double ll[5], ul[5];
  limit_bounds(r, ll, ul);
  const long n = (long)z->mx * z->my * z->mz;
  double *rep = (double *)malloc(sizeof(double) * 5 * (size_t)n);
  char *bad = (char *)calloc((size_t)n, 1);
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long q = DQX(z, i, j, k), p = IX(L, i, j, k);
        if (!point_bad(z->bv, p, sz, ll, ul)) continue;
        bad[q] = 1;
        double acc[5] = {0, 0, 0, 0, 0};
        int kn = 0;
        for (int ka = -1; ka < 2; ++ka) {
          const int k1 = k + ka;
          if (k1 < 0 || k1 >= z->mz) continue;
          for (int ja = -1; ja < 2; ++ja) {
            const int j1 = j + ja;
            if (j1 < 0 || j1 >= z->my) continue;
            for (int ia = -1; ia < 2; ++ia) {
              const int i1 = i + ia;
              if (i1 < 0 || i1 >= z->mx) continue;
              const long p1 = IX(L, i1, j1, k1);
              if (point_bad(z->bv, p1, sz, ll, ul)) continue;
              for (int l = 0; l < 5; ++l) acc[l] += z->bv[p1 + l * sz];
              ++kn;
            }
          }
        }
        if (kn > 0) {
          const double inv = 1.0 / kn;
          for (int l = 0; l < 5; ++l) acc[l] *= inv;
        } else {
          for (int l = 0; l < 5; ++l) acc[l] = dmin(dmax(z->bv[p + l * sz], ll[l]), ul[l]);
        }
        for (int l = 0; l < 5; ++l) rep[q * 5 + l] = acc[l];
      }
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long q = DQX(z, i, j, k), p = IX(L, i, j, k);
        if (!bad[q]) continue;
        for (int l = 0; l < 5; ++l) z->bv[p + l * sz] = rep[q * 5 + l];
        double R = R_AIR;
        if (r->mixture) {
          R = 0;
          for (int l = 0; l < r->n_spec; ++l) R += z->sv[p + l * sz] * tb->gas_const[l];
        }
        z->bv[p + 5 * sz] = rep[q * 5 + 4] / (rep[q * 5] * R);
        cv_from_bv_point(z, &L, r, tb, p);
      }
  free(rep);
  free(bad);
}

/* eliminate_k_gradient, src/FieldOperation.cu:96-121 */
void oc_eliminate_k_gradient(oc_zone *z, const oc_run *r) {
  const lay_t L = layout(z);
  const int g = z->ngg;
  for (int j = -g; j < z->my + g; ++j)
    for (int i = -g; i < z->mx + g; ++i) {
      const long p0 = IX(L, i, j, 0);
      for (int k = 1; k <= g; ++k) {
        const long pp = IX(L, i, j, k), pm = IX(L, i, j, -k);
        for (int l = 0; l < 6; ++l) {
          z->bv[pp + l * L.sz] = z->bv[p0 + l * L.sz];
          z->bv[pm + l * L.sz] = z->bv[p0 + l * L.sz];
        }
        for (int l = 0; l < r->n_scalar; ++l) {
          z->sv[pp + l * L.sz] = z->sv[p0 + l * L.sz];
          z->sv[pm + l * L.sz] = z->sv[p0 + l * L.sz];
        }
        for (int l = 0; l < r->n_var; ++l) {
          z->cv[pp + l * L.sz] = z->cv[p0 + l * L.sz];
          z->cv[pm + l * L.sz] = z->cv[p0 + l * L.sz];
        }
      }
    }
}

/* local_time_step (1T) + global min, src/TimeAdvanceFunc.cuh:28-127, src/TimeAdvanceFunc.cu:114-164 */
double oc_local_time_step(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  const lay_t L = layout(z);
  const long sz = L.sz;
  const int dim = z->mz == 1 ? 2 : 3;
  double dtmin = 1e+6;
#pragma omp parallel for schedule(static) reduction(min : dtmin)
  for (int k = 0; k < z->mz; ++k)
    for (int j = 0; j < z->my; ++j)
      for (int i = 0; i < z->mx; ++i) {
        const long p = IX(L, i, j, k);
        const double *mt = z->metric;
        const double grad_xi = norm3(mt[p], mt[p + sz], mt[p + 2 * sz]);
        const double grad_eta = norm3(mt[p + 3 * sz], mt[p + 4 * sz], mt[p + 5 * sz]);
        const double grad_zeta = norm3(mt[p + 6 * sz], mt[p + 7 * sz], mt[p + 8 * sz]);
        const double u = z->bv[p + sz], v = z->bv[p + 2 * sz], w = z->bv[p + 3 * sz];
        const double U = u * mt[p] + v * mt[p + sz] + w * mt[p + 2 * sz];
        const double V = u * mt[p + 3 * sz] + v * mt[p + 4 * sz] + w * mt[p + 5 * sz];
        const double W = u * mt[p + 6 * sz] + v * mt[p + 7 * sz] + w * mt[p + 8 * sz];
        const double c = r->mixture ? z->acoustic[p] : sqrt(gamma_air * R_AIR * z->bv[p + 5 * sz]);
        double msr = dmax(fabs(U) + c * grad_xi, fabs(V) + c * grad_eta);
        if (dim == 3) msr = dmax(msr, fabs(W) + c * grad_zeta);
        if (r->viscous_order > 0) {
          double max_length = dmax(grad_xi, grad_eta);
          if (dim == 3) max_length = dmax(max_length, grad_zeta);
          double mdv;
          const double iRho = 1.0 / z->bv[p];
          if (!r->mixture) {
            mdv = dmax(gamma_air, 4.0 / 3.0) * iRho;
            mdv *= z->mul[p] / r->prandtl;
          } else {
            double cpl[OC_MAX_SPEC];
            oc_cp(tb, z->bv[p + 5 * sz], cpl);
            double cp = 0, max_rhoD = 0;
            for (int l = 0; l < r->n_spec; ++l) {
              cp += cpl[l] * z->sv[p + l * sz];
              max_rhoD = dmax(max_rhoD, z->rhoD[p + l * sz]);
            }
            mdv = z->mul[p] * iRho;
            mdv = dmax(mdv, z->kappa[p] * iRho * z->gamma[p] / cp);
            mdv = dmax(mdv, max_rhoD * iRho);
            if (r->two_temperature) {
              /* src/TimeAdvanceFunc.cu:75-85 */
              double y[OC_MAX_SPEC], cv_ve = 0;
              const double tve = z->tve[p] > 0.0 ? z->tve[p] : dmax(z->bv[p + 5 * sz], 1.0);
              for (int l = 0; l < r->n_spec; ++l) y[l] = z->sv[p + l * sz];
              oc_mixture_ve_energy(tb, tve, y, &cv_ve);
              mdv = dmax(mdv, z->kappa_ve[p] * iRho / dmax(cv_ve, 1e-8));
            }
          }
          msr = dmax(msr, max_length * max_length * mdv);
        }
        double dt = r->cfl / msr;
        if (r->mixture && r->reaction) {
          const double reaction_dt = r->cfl * z->rts[p];
          if (reaction_dt > 1e-20 && reaction_dt < 1e-5) dt = dmin(dt, reaction_dt);
        }
        z->dt_local[DQX(z, i, j, k)] = dt;
        dtmin = dmin(dtmin, dt);
      }
  return dtmin;
}

/* ------------------------------------------------------------------------ */
/* orchestration                                                             */
/* ------------------------------------------------------------------------ */
/* Driver::initialize_computation, src/Driver.cu:126-211, then the RK3
 * preamble compute_cv_from_bv, src/RK.cuh:47-58 */
void oc_initialize_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb) {
  for (int b = 0; b < n_block; ++b)
    if (zones[b]->mz == 1) oc_eliminate_k_gradient(zones[b], r);
  for (int b = 0; b < n_block; ++b) oc_update_properties(zones[b], r, tb);
  for (int b = 0; b < n_block; ++b) oc_boundary_conditions(zones[b], r, tb);
  oc_data_communication(zones, n_block, r, tb);
  for (int b = 0; b < n_block; ++b) oc_update_properties(zones[b], r, tb);
  for (int b = 0; b < n_block; ++b) oc_cv_from_bv(zones[b], r, tb);
}

/* One RK3 stage, src/RK.cuh:164-236 */
void oc_stage_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb, double dt, int rk) {
  for (int b = 0; b < n_block; ++b) oc_zero_dq(zones[b], r);
  if (r->viscous_order == 8) oc_viscous8_blocks(zones, n_block, r, tb);
  if (r->viscous_order == 2)
    for (int b = 0; b < n_block; ++b) oc_viscous2(zones[b], r, tb);
  for (int b = 0; b < n_block; ++b) {
    oc_zone *z = zones[b];
    if (r->mixture && r->reaction) oc_chemistry(z, r, tb);
    oc_inviscid_dir(z, r, tb, 0);
    oc_inviscid_dir(z, r, tb, 1);
    if (z->mz > 1) oc_inviscid_dir(z, r, tb, 2);
    oc_update_rk(z, r, tb, dt, rk);
    if (r->two_temperature) oc_vt_relaxation(z, r, tb, rk_c[rk] * dt); /* src/RK.cuh:91, 200-203 */
    if (r->limit_flow) oc_limit_flow(z, r, tb);
    oc_boundary_conditions(z, r, tb);
  }
  oc_data_communication(zones, n_block, r, tb);
  for (int b = 0; b < n_block; ++b)
    if (zones[b]->mz == 1) oc_eliminate_k_gradient(zones[b], r);
  for (int b = 0; b < n_block; ++b) oc_update_properties(zones[b], r, tb);
}

/* One RK3 step, src/RK.cuh:107-272; the time step is the minimum over blocks
 * (global_time_step, src/TimeAdvanceFunc.cu:114-139).  Returns the dt used. */
double oc_step_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb, double dt_fixed,
                      int fixed) {
  for (int b = 0; b < n_block; ++b) {
    const lay_t L = layout(zones[b]);
    memcpy(zones[b]->qn, zones[b]->cv, sizeof(double) * L.sz * r->n_var);
  }
  /* compute_shock_sensor once per step for the hybrid WENO (src/RK.cuh:138-143) */
  const int sch = r->inviscid_scheme;
  if ((sch == 51 || sch == 52 || sch == 71 || sch == 72) && r->sensor_threshold > 1e-10)
    for (int b = 0; b < n_block; ++b) oc_shock_sensor(zones[b], r);
  double dt = dt_fixed;
  if (!fixed) {
    dt = 1e+6;
    for (int b = 0; b < n_block; ++b) dt = dmin(dt, oc_local_time_step(zones[b], r, tb));
  }
  for (int rk = 0; rk < 3; ++rk) oc_stage_blocks(zones, n_block, r, tb, dt, rk);
  return dt;
}

/* ------------------------------------------------------------------------ */
/* Wu's operator splitting, src/StrangSplitting.cu:15-865 (one temperature)   */
/* ------------------------------------------------------------------------ */
void oc_compute_cn_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb) {
  const int sch = r->inviscid_scheme;
  if ((sch == 51 || sch == 52 || sch == 71 || sch == 72) && r->sensor_threshold > 1e-10)
    for (int b = 0; b < n_block; ++b) oc_shock_sensor(zones[b], r);
  for (int b = 0; b < n_block; ++b) oc_zero_dq(zones[b], r);
  if (r->viscous_order == 8) oc_viscous8_blocks(zones, n_block, r, tb);
  if (r->viscous_order == 2)
    for (int b = 0; b < n_block; ++b) oc_viscous2(zones[b], r, tb);
  for (int b = 0; b < n_block; ++b) {
    oc_zone *z = zones[b];
    oc_inviscid_dir(z, r, tb, 0);
    oc_inviscid_dir(z, r, tb, 1);
    if (z->mz > 1) oc_inviscid_dir(z, r, tb, 2);
    memcpy(z->cn, z->dq, sizeof(double) * DQSZ(z) * r->n_var);
  }
}

/* wu_zeroDReaction, src/StrangSplitting.cu:43-228: over dt the momentum and
 * energy advance by the frozen transport part Cn, the partial densities by
 * the chemical source plus their Cn; chem_advance_method 3 takes three
 * SSP-RK3 stages (rates re-evaluated at each stage's T), 1 divides the source
 * by I - dt J with the mass-action Jacobian, 0 is explicit; the chemical time
 * scale min |rho Y / omega_d| goes to reaction_timeScale. */
static void lu_decomp(double *a, int dim, int *ipiv);
static void lu_solve(const double *a, double *b, int dim, const int *ipiv);

static void wu_zero_d(oc_zone *z, const lay_t *L, const oc_run *r, const oc_tables *tb, long p, double dt,
                      const double *cn, const double *cnY) {
  const int ns = tb->n_spec, nr = tb->n_reac;
  const long sz = L->sz;
  double *cv = z->cv, *bv = z->bv;
  double qn[5 + OC_MAX_SPEC], yc[OC_MAX_SPEC] = {0}, omega[OC_MAX_SPEC], omega_d[OC_MAX_SPEC];
  double q1[OC_MAX_REAC], q2[OC_MAX_REAC];
  for (int l = 0; l < 5 + ns; ++l) qn[l] = cv[p + l * sz];
  const double iJac = 1.0 / z->jac[p];
  if (r->chem_advance_method == 3) {
    for (int stage = 0; stage < 3; ++stage) {
      const double t_stage = bv[p + 5 * sz];
      /* concentrations c = rho Y imw 1e-3 from the partial densities (rho = 1 times rho Y) */
      for (int l = 0; l < ns; ++l) yc[l] = cv[p + (5 + l) * sz];
      chem_rates(tb, 1.0, t_stage, t_stage, yc, omega, omega_d, q1, q2);
      for (int l = 0; l < 5; ++l)
        cv[p + l * sz] = rk_a[stage] * qn[l] + rk_b[stage] * cv[p + l * sz] + rk_c[stage] * dt * iJac * cn[l];
      double time_scale = 1e+6;
      for (int l = 0; l < ns; ++l) {
        double *c = &cv[p + (5 + l) * sz];
        if (*c > 1e-25 && fabs(omega_d[l]) > 1e-25) time_scale = dmin(time_scale, fabs(*c / omega_d[l]));
        *c = rk_a[stage] * qn[5 + l] + rk_b[stage] * *c + rk_c[stage] * dt * (omega[l] + cnY[l] * iJac);
      }
      z->rts[p] = time_scale;
      recover_point(z, L, r, tb, p);
    }
    return;
  }
  const double t = bv[p + 5 * sz];
  for (int l = 0; l < ns; ++l) yc[l] = cv[p + (5 + l) * sz];
  chem_rates(tb, 1.0, t, t, yc, omega, omega_d, q1, q2);
  if (r->chem_advance_method == 1) {
    double c[OC_MAX_SPEC], lhs[OC_MAX_SPEC * OC_MAX_SPEC];
    int ipiv[OC_MAX_SPEC];
    for (int l = 0; l < ns; ++l) c[l] = yc[l] * tb->imw[l] * 1e-3 * (1e+3 / tb->imw[l]);
    for (int m = 0; m < ns; ++m)
      for (int n = 0; n < ns; ++n) {
        double zz = 0;
        if (c[n] > 1e-25) {
          for (int i = 0; i < nr; ++i)
            zz += (tb->stoi_b[i][m] - tb->stoi_f[i][m]) * (tb->stoi_f[i][n] * q1[i] - tb->stoi_b[i][n] * q2[i]);
          zz /= c[n];
        }
        lhs[m * ns + n] = (m == n ? 1.0 : 0.0) - dt * (zz * 1e+3 / tb->imw[m]);
      }
    lu_decomp(lhs, ns, ipiv);
    lu_solve(lhs, omega, ns, ipiv);
  }
  for (int l = 0; l < 5; ++l) cv[p + l * sz] += cn[l] * dt * iJac;
  double time_scale = 1e+6;
  for (int l = 0; l < ns; ++l) {
    double *cl = &cv[p + (5 + l) * sz];
    if (*cl > 1e-25 && fabs(omega_d[l]) > 1e-25) time_scale = dmin(time_scale, fabs(*cl / omega_d[l]));
    *cl += (omega[l] + cnY[l] * iJac) * dt;
  }
  z->rts[p] = time_scale;
  recover_point(z, L, r, tb, p);
}

void oc_wu_reaction(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt) {
  const lay_t L = layout(z);
  const int ns = tb->n_spec;
  const long dsz = DQSZ(z);
#pragma omp parallel for schedule(dynamic, 64)
  for (long q = 0; q < dsz; ++q) {
    const int i = (int)(q % z->mx), j = (int)((q / z->mx) % z->my), k = (int)(q / ((long)z->mx * z->my));
    const long p = IX(L, i, j, k);
    double cn[5], cnY[OC_MAX_SPEC];
    for (int l = 0; l < 5; ++l) cn[l] = z->cn[q + l * dsz];
    for (int l = 0; l < ns; ++l) cnY[l] = z->cn[q + (5 + l) * dsz];
    double ts = z->rts[p];
    if (ts < 1e-15) ts = 0.2 * dt;
    const int n_sub = (int)(dt / ts);
    double dt_reaction = dt / (n_sub + 1);
    for (int n = 0; n < n_sub; ++n) wu_zero_d(z, &L, r, tb, p, dt_reaction, cn, cnY);
    dt_reaction = dt - dt_reaction * n_sub;
    if (fabs(dt_reaction) > 1e-12) wu_zero_d(z, &L, r, tb, p, dt_reaction, cn, cnY);
  }
}

double oc_wu_step_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb, double dt_fixed,
                         int fixed) {
  double dt = dt_fixed;
  if (!fixed) {
    /* local_time_step_without_reaction, src/TimeAdvanceFunc.cu:21-112 */
    oc_run rr = *r;
    rr.reaction = 0;
    dt = 1e+6;
    for (int b = 0; b < n_block; ++b) dt = dmin(dt, oc_local_time_step(zones[b], &rr, tb));
  }
  oc_compute_cn_blocks(zones, n_block, r, tb);
  /* wu_reaction_subStep, src/StrangSplitting.cu:479-506 */
  for (int b = 0; b < n_block; ++b) {
    if (r->reaction) oc_wu_reaction(zones[b], r, tb, dt);
    oc_boundary_conditions(zones[b], r, tb);
  }
  oc_data_communication(zones, n_block, r, tb);
  for (int b = 0; b < n_block; ++b)
    if (zones[b]->mz == 1) oc_eliminate_k_gradient(zones[b], r);
  for (int b = 0; b < n_block; ++b) oc_update_properties(zones[b], r, tb);
  /* wu_RK_subStep over dt / 2, src/StrangSplitting.cu:568-638 (no property update inside) */
  const double dh = 0.5 * dt;
  for (int b = 0; b < n_block; ++b) {
    const lay_t L = layout(zones[b]);
    memcpy(zones[b]->qn, zones[b]->cv, sizeof(double) * L.sz * r->n_var);
  }
  for (int rk = 0; rk < 3; ++rk) {
    for (int b = 0; b < n_block; ++b) oc_zero_dq(zones[b], r);
    if (r->viscous_order == 8) oc_viscous8_blocks(zones, n_block, r, tb);
    if (r->viscous_order == 2)
      for (int b = 0; b < n_block; ++b) oc_viscous2(zones[b], r, tb);
    for (int b = 0; b < n_block; ++b) {
      oc_zone *z = zones[b];
      oc_inviscid_dir(z, r, tb, 0);
      oc_inviscid_dir(z, r, tb, 1);
      if (z->mz > 1) oc_inviscid_dir(z, r, tb, 2);
      oc_update_rk_wu(z, r, tb, dh, rk);
      if (r->limit_flow) oc_limit_flow(z, r, tb);
      oc_boundary_conditions(z, r, tb);
    }
    oc_data_communication(zones, n_block, r, tb);
    for (int b = 0; b < n_block; ++b)
      if (zones[b]->mz == 1) oc_eliminate_k_gradient(zones[b], r);
  }
  return dt;
}

void oc_initialize(oc_zone *z, const oc_run *r, const oc_tables *tb) {
  oc_zone *zs[1] = {z};
  oc_initialize_blocks(zs, 1, r, tb);
}

void oc_stage(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk) {
  oc_zone *zs[1] = {z};
  oc_stage_blocks(zs, 1, r, tb, dt, rk);
}

double oc_step(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt_fixed, int fixed) {
  oc_zone *zs[1] = {z};
  return oc_step_blocks(zs, 1, r, tb, dt_fixed, fixed);
}

/* ------------------------------------------------------------------------ */
/* 0-D constant-volume reactor (RK3 branch), src/ConstVolumeReactor.cpp      */
/* ------------------------------------------------------------------------ */
/* update_t, src/ConstVolumeReactor.cpp:1456-1482 */
static double update_t(const oc_tables *tb, double T0, double imw, const double *yk, double E) {
  double t = T0, err = 1e+6;
  const double R = R_u * imw;
  int iter = 0;
  double hk[OC_MAX_SPEC], cpk[OC_MAX_SPEC];
  while (err > 1e-8 && iter++ < 100) {
    oc_enthalpy_and_cp(tb, t, hk, cpk);
    double cp_tot = 0, h = 0;
    for (int l = 0; l < tb->n_spec; ++l) {
      cp_tot += cpk[l] * yk[l];
      h += hk[l] * yk[l];
    }
    const double et = h - R * t;
    const double cv = cp_tot - R;
    const double t1 = t - (et - E) / cv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

long oc_cv_reactor(const oc_tables *tb, double rho, double *t_inout, double *y_inout, double dt, double t_end,
                   double *trace_t, double *trace_T, long trace_every, long trace_cap) {
  const int ns = tb->n_spec;
  double yk[OC_MAX_SPEC], rhoY[OC_MAX_SPEC], rhoYN[OC_MAX_SPEC], omega[OC_MAX_SPEC], omega_d[OC_MAX_SPEC];
  double T = *t_inout, hk[OC_MAX_SPEC], imw = 0;
  for (int l = 0; l < ns; ++l) {
    yk[l] = y_inout[l];
    rhoY[l] = rho * yk[l];
    imw += yk[l] * tb->imw[l];
  }
  /* E = sum Y h - p/rho = sum Y h - R T, src/ConstVolumeReactor.cpp:324-334 */
  oc_enthalpy(tb, T, hk);
  double E = 0;
  for (int l = 0; l < ns; ++l) E += yk[l] * hk[l];
  E -= R_u * imw * T;
  double time = 0;
  long step = 0, nt = 0;
  oc_chem_source(tb, rho, T, T, yk, omega, omega_d);
  while (time < t_end - 1e-30) {
    double tmp_y[OC_MAX_SPEC];
    /* stage 1 */
    imw = 0;
    for (int l = 0; l < ns; ++l) {
      rhoYN[l] = rhoY[l] + omega[l] * dt;
      tmp_y[l] = rhoYN[l] / rho;
      imw += tmp_y[l] * tb->imw[l];
    }
    double Ts = update_t(tb, T, imw, tmp_y, E);
    oc_chem_source(tb, rho, Ts, Ts, tmp_y, omega, omega_d);
    /* stage 2 */
    imw = 0;
    for (int l = 0; l < ns; ++l) {
      rhoYN[l] = 0.75 * rhoY[l] + 0.25 * (rhoYN[l] + omega[l] * dt);
      tmp_y[l] = rhoYN[l] / rho;
      imw += tmp_y[l] * tb->imw[l];
    }
    Ts = update_t(tb, Ts, imw, tmp_y, E);
    oc_chem_source(tb, rho, Ts, Ts, tmp_y, omega, omega_d);
    /* stage 3 */
    imw = 0;
    for (int l = 0; l < ns; ++l) {
      rhoY[l] = (rhoY[l] + 2.0 * (rhoYN[l] + omega[l] * dt)) / 3.0;
      if (rhoY[l] < 0) rhoY[l] = 0;
      yk[l] = rhoY[l] / rho;
      imw += yk[l] * tb->imw[l];
    }
    T = update_t(tb, T, imw, yk, E);
    oc_chem_source(tb, rho, T, T, yk, omega, omega_d);
    time += dt;
    ++step;
    if (trace_t && trace_every > 0 && step % trace_every == 0 && nt < trace_cap) {
      trace_t[nt] = time;
      trace_T[nt] = T;
      ++nt;
    }
  }
  *t_inout = T;
  for (int l = 0; l < ns; ++l) y_inout[l] = yk[l];
  return step;
}

/* 0-D constant-volume reactor with the point-implicit update (advance_scheme 0,
 * implicit_method 2 "DA", src/ConstVolumeReactor.cpp:400-440): the source is
 * divided by 1 - dt J_ll with the diagonal of the source Jacobian
 * (compute_chem_src_jacobian_diagonal, 1760-1777; DA, 1795-1799), rho Y
 * advances by the modified source and is clipped at 0 (with conserve_mass
 * the mass fractions are renormalised, 418-430), then T follows from update_t
 * (1456-1482); time advances by the fixed dt until it reaches t_end.
 * scheme 3 is the explicit SSP-RK3 of oc_cv_reactor. */
/* compute_chem_src_jacobian, src/ConstVolumeReactor.cpp:1739-1758:
 * J[m][n] = 1e3 MW_m / (rho Y_n) sum_r (nu_b - nu_f)_rm (nu_f,rn q1_r - nu_b,rn q2_r), 0 where rho Y_n <= 1e-30 */
static void chem_jacobian(const oc_tables *tb, const double *rhoY, const double *q1, const double *q2, double *jac) {
  const int ns = tb->n_spec, nr = tb->n_reac;
  for (int m = 0; m < ns; ++m)
    for (int n = 0; n < ns; ++n) {
      double zz = 0;
      if (rhoY[n] > 1e-30) {
        for (int r = 0; r < nr; ++r)
          zz += (tb->stoi_b[r][m] - tb->stoi_f[r][m]) * (tb->stoi_f[r][n] * q1[r] - tb->stoi_b[r][n] * q2[r]);
        zz /= rhoY[n];
      }
      jac[m * ns + n] = zz * 1e+3 / tb->imw[m];
    }
}

/* lu_decomp (column-pivoted, row swaps recorded in ipiv) and lu_to_solution,
 * src/ConstVolumeReactor.cpp:1802-1856, in the same operation order */
static void lu_decomp(double *a, int dim, int *ipiv) {
  for (int n = 0; n < dim; ++n) {
    int ik = n;
    for (int m = n; m < dim; ++m) {
      for (int t = 0; t < n; ++t) a[m * dim + n] -= a[m * dim + t] * a[t * dim + n];
      if (fabs(a[m * dim + n]) > fabs(a[ik * dim + n])) ik = m;
    }
    ipiv[n] = ik;
    if (ik != n)
      for (int t = 0; t < dim; ++t) {
        const double mid = a[ik * dim + t];
        a[ik * dim + t] = a[n * dim + t];
        a[n * dim + t] = mid;
      }
    for (int q = n + 1; q < dim; ++q)
      for (int t = 0; t < n; ++t) a[n * dim + q] -= a[n * dim + t] * a[t * dim + q];
    for (int m = n + 1; m < dim; ++m) a[m * dim + n] /= a[n * dim + n];
  }
}

static void lu_solve(const double *a, double *b, int dim, const int *ipiv) {
  for (int m = 0; m < dim; ++m) {
    const int t = ipiv[m];
    if (t != m) {
      const double mid = b[t];
      b[t] = b[m];
      b[m] = mid;
    }
  }
  for (int m = 1; m < dim; ++m)
    for (int t = 0; t < m; ++t) b[m] -= a[m * dim + t] * b[t];
  b[dim - 1] /= a[dim * dim - 1];
  for (int m = dim - 2; m >= 0; --m) {
    for (int t = m + 1; t < dim; ++t) b[m] -= a[m * dim + t] * b[t];
    b[m] /= a[m * dim + m];
  }
}

/* ROK4E coefficients (the method's published tableau, as src/Define.h:33-48) */
static const double rk_gamma = 0.572816062482135;
static const double rk_a21 = 0.432364435748567, rk_a31 = -0.51421131687617, rk_a32 = 1.38227114461736;
#define RK_G21 (-0.602765307997356 / rk_gamma)
#define RK_G31 (-1.389195789724843 / rk_gamma)
#define RK_G32 (1.072950969011413 / rk_gamma)
#define RK_G41 (0.992356412977094 / rk_gamma)
#define RK_G42 (-1.390032613873701 / rk_gamma)
#define RK_G43 (-0.440875890223325 / rk_gamma)
static const double rk_b1 = 0.194335256262729, rk_b2 = 0.483167813989227, rk_b4 = 0.322496929748044;
#define RK_E1 (-0.217819895945721 - rk_b1)
#define RK_E2 (1.03130847478467 - rk_b2)
#define RK_E3 0.186511421161047
#define RK_E4 (-rk_b4)

/* ROK4E with the full Jacobian (advance_scheme 1, build_arnoldi 0,
 * src/ConstVolumeReactor.cpp:442-707 and 748-762): four stages with
 * LHS = I - h gamma J factored once; with adaptive steps the embedded error
 * sqrt(mean((h e.k / (RTol rhoY_new + ATol))^2)) accepts (<= 1) or rejects the
 * step and scales h by min(5, max(0.2, 0.8 err_last^0.1 (err + 1e-20)^-0.175))
 * (0.2 on a rejected first step); h never passes t_end.  Returns accepted steps;
 * *rejected_out counts rejections. */
long oc_cv_reactor_rok4e(const oc_tables *tb, double rho, double *t_inout, double *y_inout, double h0, double t_end,
                         int adaptive, double rtol, double atol, long *rejected_out) {
  const int ns = tb->n_spec;
  double yk[OC_MAX_SPEC], rhoY[OC_MAX_SPEC], omega[OC_MAX_SPEC], omega_d[OC_MAX_SPEC], hk[OC_MAX_SPEC];
  double q1[OC_MAX_REAC], q2[OC_MAX_REAC], jac[OC_MAX_SPEC * OC_MAX_SPEC], lhs[OC_MAX_SPEC * OC_MAX_SPEC];
  double k1[OC_MAX_SPEC], k2[OC_MAX_SPEC], k3[OC_MAX_SPEC], k4[OC_MAX_SPEC], ys[OC_MAX_SPEC];
  int ipiv[OC_MAX_SPEC];
  double T = *t_inout, imw = 0;
  for (int l = 0; l < ns; ++l) {
    yk[l] = y_inout[l];
    rhoY[l] = rho * yk[l];
    imw += yk[l] * tb->imw[l];
  }
  oc_enthalpy(tb, T, hk);
  double E = 0;
  for (int l = 0; l < ns; ++l) E += yk[l] * hk[l];
  E -= R_u * imw * T;
  double time = 0, h = h0, err_last = 0;
  long step = 0, rejected = 0;
  while (time < t_end) {
    chem_rates(tb, rho, T, T, yk, omega, omega_d, q1, q2);
    const double hs = dmin(h, t_end - time);
    chem_jacobian(tb, rhoY, q1, q2, jac);
    for (int m = 0; m < ns; ++m)
      for (int n = 0; n < ns; ++n) lhs[m * ns + n] = (m == n ? 1.0 : 0.0) - hs * rk_gamma * jac[m * ns + n];
    lu_decomp(lhs, ns, ipiv);
    for (int l = 0; l < ns; ++l) k1[l] = omega[l];
    lu_solve(lhs, k1, ns, ipiv);
    /* stage 2 */
    imw = 0;
    for (int l = 0; l < ns; ++l) {
      ys[l] = rhoY[l] + hs * rk_a21 * k1[l];
      yk[l] = ys[l] / rho;
      imw += yk[l] * tb->imw[l];
    }
    double Ts = update_t(tb, T, imw, yk, E);
    chem_rates(tb, rho, Ts, Ts, yk, omega, omega_d, q1, q2);
    for (int l = 0; l < ns; ++l) k2[l] = omega[l] + RK_G21 * k1[l];
    lu_solve(lhs, k2, ns, ipiv);
    for (int l = 0; l < ns; ++l) k2[l] -= RK_G21 * k1[l];
    /* stage 3 */
    imw = 0;
    for (int l = 0; l < ns; ++l) {
      ys[l] = rhoY[l] + hs * (rk_a31 * k1[l] + rk_a32 * k2[l]);
      yk[l] = ys[l] / rho;
      imw += yk[l] * tb->imw[l];
    }
    Ts = update_t(tb, Ts, imw, yk, E);
    chem_rates(tb, rho, Ts, Ts, yk, omega, omega_d, q1, q2);
    for (int l = 0; l < ns; ++l) k3[l] = omega[l] + RK_G31 * k1[l] + RK_G32 * k2[l];
    lu_solve(lhs, k3, ns, ipiv);
    for (int l = 0; l < ns; ++l) k3[l] -= RK_G31 * k1[l] + RK_G32 * k2[l];
    /* stage 4 (the source of stage 3) */
    for (int l = 0; l < ns; ++l) k4[l] = omega[l] + RK_G41 * k1[l] + RK_G42 * k2[l] + RK_G43 * k3[l];
    lu_solve(lhs, k4, ns, ipiv);
    for (int l = 0; l < ns; ++l) k4[l] -= RK_G41 * k1[l] + RK_G42 * k2[l] + RK_G43 * k3[l];
    int accepted = 1;
    if (adaptive) {
      double rn[OC_MAX_SPEC], err = 0;
      for (int l = 0; l < ns; ++l) {
        rn[l] = rhoY[l] + hs * (rk_b1 * k1[l] + rk_b2 * k2[l] + rk_b4 * k4[l]);
        double dy = hs * (RK_E1 * k1[l] + RK_E2 * k2[l] + RK_E3 * k3[l] + RK_E4 * k4[l]);
        dy /= rtol * rn[l] + atol;
        err += dy * dy;
      }
      err = sqrt(err / ns);
      double hstar = hs;
      if (step > 0)
        hstar *= dmin(5.0, dmax(0.2, 0.8 * pow(err_last, 0.1) * pow(err + 1e-20, -0.175)));
      else if (err > 1)
        hstar *= 0.2;
      if (err <= 1) {
        for (int l = 0; l < ns; ++l) {
          rhoY[l] = rn[l];
          yk[l] = rn[l] / rho;
        }
        time += hs;
        err_last = err;
      } else {
        accepted = 0;
        ++rejected;
        /* the stage states leave yk; the next attempt starts from rhoY */
        for (int l = 0; l < ns; ++l) yk[l] = rhoY[l] / rho;
      }
      h = hstar;
    } else {
      for (int l = 0; l < ns; ++l) {
        rhoY[l] += hs * (rk_b1 * k1[l] + rk_b2 * k2[l] + rk_b4 * k4[l]);
        yk[l] = rhoY[l] / rho;
      }
      time += hs;
    }
    if (!accepted) continue;
    imw = 0;
    for (int l = 0; l < ns; ++l) imw += yk[l] * tb->imw[l];
    T = update_t(tb, T, imw, yk, E);
    ++step;
  }
  *t_inout = T;
  for (int l = 0; l < ns; ++l) y_inout[l] = yk[l];
  if (rejected_out) *rejected_out = rejected;
  return step;
}

long oc_cv_reactor_ex(const oc_tables *tb, double rho, double *t_inout, double *y_inout, double dt, double t_end,
                      int scheme, int conserve_mass) {
  if (scheme == 3) return oc_cv_reactor(tb, rho, t_inout, y_inout, dt, t_end, NULL, NULL, 0, 0);
  const int ns = tb->n_spec, nr = tb->n_reac;
  double yk[OC_MAX_SPEC], rhoY[OC_MAX_SPEC], omega[OC_MAX_SPEC], omega_d[OC_MAX_SPEC], hk[OC_MAX_SPEC];
  double q1[OC_MAX_REAC], q2[OC_MAX_REAC];
  double T = *t_inout, imw = 0;
  for (int l = 0; l < ns; ++l) {
    yk[l] = y_inout[l];
    rhoY[l] = rho * yk[l];
    imw += yk[l] * tb->imw[l];
  }
  oc_enthalpy(tb, T, hk);
  double E = 0;
  for (int l = 0; l < ns; ++l) E += yk[l] * hk[l];
  E -= R_u * imw * T;
  double time = 0;
  long step = 0;
  while (time < t_end) {
    chem_rates(tb, rho, T, T, yk, omega, omega_d, q1, q2);
    for (int m = 0; m < ns; ++m) {
      double zz = 0;
      if (rhoY[m] > 1e-30) {
        for (int r = 0; r < nr; ++r)
          zz += (tb->stoi_b[r][m] - tb->stoi_f[r][m]) * (tb->stoi_f[r][m] * q1[r] - tb->stoi_b[r][m] * q2[r]);
        zz /= rhoY[m];
      }
      const double jac = zz * 1e+3 / tb->imw[m];
      omega[m] /= 1 - dt * jac;
    }
    if (conserve_mass) {
      double sum = 0;
      for (int l = 0; l < ns; ++l) {
        rhoY[l] += omega[l] * dt;
        if (rhoY[l] < 0) rhoY[l] = 0;
        sum += rhoY[l];
      }
      for (int l = 0; l < ns; ++l) {
        yk[l] = rhoY[l] / sum;
        rhoY[l] = yk[l] * rho;
      }
    } else {
      for (int l = 0; l < ns; ++l) {
        rhoY[l] += omega[l] * dt;
        if (rhoY[l] < 0) rhoY[l] = 0;
      }
      for (int l = 0; l < ns; ++l) yk[l] = rhoY[l] / rho;
    }
    imw = 0;
    for (int l = 0; l < ns; ++l) imw += yk[l] * tb->imw[l];
    T = update_t(tb, T, imw, yk, E);
    time += dt;
    ++step;
  }
  *t_inout = T;
  for (int l = 0; l < ns; ++l) y_inout[l] = yk[l];
  return step;
}
