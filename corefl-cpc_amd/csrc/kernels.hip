// gfx950 kernels for COREFL's per-stage right-hand side and SSP-RK3 update.
//
// Layout: the reference's DZone planes (i fastest, ghosts included).  All
// arithmetic is FP64.  Per-species arrays are sized by the template parameter
// NS so that they live in VGPRs; NS == 0 selects the Air perfect gas.
#include <hip/hip_runtime.h>

#include "corefl_internal.h"
#include "fastmath.h"

namespace corefl {
namespace {

#define DEV __device__ __forceinline__

constexpr double kPatm = 101325;
constexpr double kGammaAir = 1.4;
constexpr double kMwAir = 28.9647;
constexpr double kRair = kRu / kMwAir;
constexpr double kPi = 3.141592653589793;
constexpr double kEpsWeno = 1e-40;  // src/WENO.cu:8

// SSP-RK3 coefficients, src/RK.cuh:12-16
__constant__ double c_rk_a[3] = {1.0, 0.75, 1.0 / 3.0};
__constant__ double c_rk_b[3] = {0.0, 0.25, 2.0 / 3.0};
__constant__ double c_rk_c[3] = {1.0, 0.25, 2.0 / 3.0};

constexpr int kBlock = 256;
constexpr int kMaxNv = 5 + COREFL_MAX_SPEC + 1;  // conservative variables: 5, species, E_ve

// Thermochemical model code of the species-templated kernels: species count in
// the low five bits, NASA-9 thermodynamics (else NASA-7) in bit 5, the
// two-temperature model (vibrational-electronic energy as scalar n_spec) in
// bit 6.  0 is the Air perfect gas.
constexpr int kNs(int m) { return m & 31; }
constexpr bool kNasa9(int m) { return (m & 32) != 0; }
constexpr bool kTT(int m) { return (m & 64) != 0; }
constexpr int kNsc(int m) { return kNs(m) + (kTT(m) ? 1 : 0); }  // n_scalar
constexpr int kNv(int m) { return 5 + kNsc(m); }                  // n_var

DEV double norm3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

// ---------------------------------------------------------------------------
// NASA-7 thermodynamics (two ranges, linear extrapolation outside), the
// Combustion2Part branch of src/Thermo.cu:386-631.
// ---------------------------------------------------------------------------
DEV double nasa_h(const double *c, double t, double t2, double t3, double t4, double t5) {
  return c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
}
DEV double nasa_cp(const double *c, double t, double t2, double t3, double t4) {
  return c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
}

// NASA-9 (src/Thermo.cu:426-456, 501-521, 573-596, 628-653): the range of
// species i at t is the first j with t_cut[j] <= t <= t_cut[j+1]; outside
// [t_cut[0], t_cut[n]] the enthalpy is extrapolated linearly with the bound's
// cp, cp is held and G/RT is taken at the bound.  Values are per unit R_i.
DEV double nasa9_h_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt, itt = 1.0 / tt;
  return -c[0] * itt + c[1] * log(tt) + c[2] * tt + 0.5 * c[3] * tt2 + c[4] * tt3 / 3 + 0.25 * c[5] * tt4 +
         0.2 * c[6] * tt5 + c[7];
}
DEV double nasa9_cp_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, itt2 = 1.0 / tt2, itt = 1.0 / tt;
  return c[0] * itt2 + c[1] * itt + c[2] + c[3] * tt + c[4] * tt2 + c[5] * tt3 + c[6] * tt4;
}
DEV double nasa9_g_at(const double *c, double tt) {
  const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, itt = 1 / tt, lntt = log(tt), itt2 = itt * itt;
  return -0.5 * c[0] * itt2 + c[1] * itt * (lntt + 1) + c[2] * (1.0 - lntt) - 0.5 * c[3] * tt - c[4] * tt2 / 6.0 -
         c[5] * tt3 / 12.0 - c[6] * tt4 * 0.05 + c[7] * itt - c[8];
}
// range index inside [t_cut[0], t_cut[n]] (the interior cuts below t)
DEV int nasa9_range(const Tab *__restrict__ tb, int i, double t) {
  const int n = tb->n_range[i];
  int j = 0;
#pragma unroll
  for (int m = 1; m < COREFL_MAX_RANGE; ++m) j += (m < n && tb->t_cut[i][m] < t) ? 1 : 0;
  return j;
}
// h_i / R_i and cp_i / R_i (either may be skipped)
template <bool WANT_H, bool WANT_CP>
DEV void nasa9_h_cp(const Tab *__restrict__ tb, int i, double t, double &h, double &cp) {
  const int n = tb->n_range[i];
  const double lo = tb->t_cut[i][0], hi = tb->t_cut[i][n];
  if (t < lo || t > hi) {
    const double tt = t < lo ? lo : hi;
    const double *c = tb->poly9[i][t < lo ? 0 : n - 1];
    const double cb = nasa9_cp_at(c, tt);
    if (WANT_H) h = nasa9_h_at(c, tt) + cb * (t - tt);
    if (WANT_CP) cp = cb;
  } else {
    const double *c = tb->poly9[i][nasa9_range(tb, i, t)];
    const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, it = 1.0 / t;
    if (WANT_H) {
      const double t5 = t4 * t, lnT = log(t);
      h = -c[0] * it + c[1] * lnT + c[2] * t + 0.5 * c[3] * t2 + c[4] * t3 / 3 + 0.25 * c[5] * t4 +
          0.2 * c[6] * t5 + c[7];
    }
    if (WANT_CP) {
      const double it2 = it * it;
      cp = c[0] * it2 + c[1] * it + c[2] + c[3] * t + c[4] * t2 + c[5] * t3 + c[6] * t4;
    }
  }
}

template <int M>
DEV void thermo_h(const Tab *__restrict__ tb, double t, double *h) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  if constexpr (kNasa9(M)) {
    double cp;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      nasa9_h_cp<true, false>(tb, i, t, h[i], cp);
      h[i] *= tb->gas_const[i];
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double tl = tb->t_low[i], th = tb->t_high[i];
    double v;
    if (t < tl || t > th) {
      const double tt = t < tl ? tl : th;
      const double *c = t < tl ? tb->low_coeff[i] : tb->high_coeff[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      v = nasa_h(c, tt, tt2, tt3, tt4, tt5);
      v += nasa_cp(c, tt, tt2, tt3, tt4) * (t - tt);
    } else {
      const bool lo = t < tb->t_mid[i];
      double c[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) c[k] = lo ? tb->low_coeff[i][k] : tb->high_coeff[i][k];
      v = nasa_h(c, t, t2, t3, t4, t5);
    }
    h[i] = v * tb->gas_const[i];
  }
}

template <int M>
DEV void thermo_cp(const Tab *__restrict__ tb, double t, double *cp) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  if constexpr (kNasa9(M)) {
    double h;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      nasa9_h_cp<false, true>(tb, i, t, h, cp[i]);
      cp[i] *= tb->gas_const[i];
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double tl = tb->t_low[i], th = tb->t_high[i];
    double v;
    if (t < tl || t > th) {
      const double tt = t < tl ? tl : th;
      const double *c = t < tl ? tb->low_coeff[i] : tb->high_coeff[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      v = nasa_cp(c, tt, tt2, tt3, tt4);
    } else {
      const bool lo = t < tb->t_mid[i];
      double c[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) c[k] = lo ? tb->low_coeff[i][k] : tb->high_coeff[i][k];
      v = nasa_cp(c, t, t2, t3, t4);
    }
    cp[i] = v * tb->gas_const[i];
  }
}

template <int M>
DEV void thermo_h_cp(const Tab *__restrict__ tb, double t, double *h, double *cp) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  if constexpr (kNasa9(M)) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      nasa9_h_cp<true, true>(tb, i, t, h[i], cp[i]);
      h[i] *= tb->gas_const[i];
      cp[i] *= tb->gas_const[i];
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double tl = tb->t_low[i], th = tb->t_high[i];
    double vh, vc;
    if (t < tl || t > th) {
      const double tt = t < tl ? tl : th;
      const double *c = t < tl ? tb->low_coeff[i] : tb->high_coeff[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      vh = nasa_h(c, tt, tt2, tt3, tt4, tt5);
      vc = nasa_cp(c, tt, tt2, tt3, tt4);
      vh += vc * (t - tt);
    } else {
      const bool lo = t < tb->t_mid[i];
      double c[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) c[k] = lo ? tb->low_coeff[i][k] : tb->high_coeff[i][k];
      vh = nasa_h(c, t, t2, t3, t4, t5);
      vc = nasa_cp(c, t, t2, t3, t4);
    }
    h[i] = vh * tb->gas_const[i];
    cp[i] = vc * tb->gas_const[i];
  }
}

// h_i(t) of one species (same arithmetic as thermo_h).
template <int M>
DEV double thermo_h1(const Tab *__restrict__ tb, int i, double t, double t2, double t3, double t4, double t5) {
  if constexpr (kNasa9(M)) {
    double h, cp;
    nasa9_h_cp<true, false>(tb, i, t, h, cp);
    return h * tb->gas_const[i];
  }
  const double tl = tb->t_low[i], th = tb->t_high[i];
  double v;
  if (t < tl || t > th) {
    const double tt = t < tl ? tl : th;
    const double *c = t < tl ? tb->low_coeff[i] : tb->high_coeff[i];
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
    v = nasa_h(c, tt, tt2, tt3, tt4, tt5);
    v += nasa_cp(c, tt, tt2, tt3, tt4) * (t - tt);
  } else {
    const bool lo = t < tb->t_mid[i];
    double c[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) c[k] = lo ? tb->low_coeff[i][k] : tb->high_coeff[i][k];
    v = nasa_h(c, t, t2, t3, t4, t5);
  }
  return v * tb->gas_const[i];
}

// compute_gibbs_div_rt, NASA-7 branch, src/Thermo.cu:600-626
template <int M>
DEV void thermo_gibbs(const Tab *__restrict__ tb, double t, double *g) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  if constexpr (kNasa9(M)) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const int n = tb->n_range[i];
      const double lo = tb->t_cut[i][0], hi = tb->t_cut[i][n];
      if (t < lo || t > hi)
        g[i] = nasa9_g_at(tb->poly9[i][t < lo ? 0 : n - 1], t < lo ? lo : hi);
      else
        g[i] = nasa9_g_at(tb->poly9[i][nasa9_range(tb, i, t)], t);
    }
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double tl = tb->t_low[i], th = tb->t_high[i];
    if (t < tl || t > th) {
      const double tt = t < tl ? tl : th;
      const double *c = t < tl ? tb->low_coeff[i] : tb->high_coeff[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
      g[i] = c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
             c[5] * tt_inv - c[6];
    } else {
      const bool lo = t < tb->t_mid[i];
      double c[7];
#pragma unroll
      for (int k = 0; k < 7; ++k) c[k] = lo ? tb->low_coeff[i][k] : tb->high_coeff[i][k];
      g[i] = c[0] * (1.0 - log_t) - 0.5 * c[1] * t - c[2] * t2 / 6.0 - c[3] * t3 / 12.0 - c[4] * t4 * 0.05 +
             c[5] * t_inv - c[6];
    }
  }
}

// ---------------------------------------------------------------------------
// Two-temperature energy modes and V-T relaxation, src/Thermo.cu:635-901 (per
// unit mass; temperatures clamped at 1 K).  Only the two-temperature models
// call these.
// ---------------------------------------------------------------------------
DEV double vib_energy1(const Tab *__restrict__ tb, int i, double t) {
  const double temp = fmax(t, 1.0);
  double e = 0.0;
  const double th = tb->theta_v[i];
  if (th > 0) {
    const double x = th / temp;
    if (x < 200) e += tb->gas_const[i] * th / expm1(x);
  }
  return e;
}
DEV double vib_cv1(const Tab *__restrict__ tb, int i, double t) {
  const double temp = fmax(t, 1.0);
  double c = 0.0;
  const double th = tb->theta_v[i];
  if (th > 0) {
    const double x = th / temp;
    if (x < 200) {
      const double ex = exp(x), d = ex - 1.0;
      c += tb->gas_const[i] * x * x * ex / (d * d);
    }
  }
  return c;
}
// E_ve of species i (compute_ve_energy) and, when WANT_CV, its c_v (compute_ve_cv)
template <bool WANT_CV>
DEV double ve_energy1(const Tab *__restrict__ tb, int i, double t, double *cv) {
  const double temp = fmax(t, 1.0);
  double e = vib_energy1(tb, i, temp);
  double c = WANT_CV ? vib_cv1(tb, i, temp) : 0.0;
  const int n = tb->n_elec[i];
  if (n > 0) {
    double zz = 0.0, b = 0.0, cc = 0.0;
    for (int j = 0; j < n; ++j) {
      const double th = tb->elec_theta[i][j], boltz = tb->elec_g[i][j] * exp(-th / temp);
      zz += boltz;
      b += boltz * th;
      if (WANT_CV) cc += boltz * th * th;
    }
    if (zz > 0) {
      e += tb->gas_const[i] * b / zz;
      if (WANT_CV) {
        const double mean = b / zz;
        c += tb->gas_const[i] * (cc / zz - mean * mean) / (temp * temp);
      }
    }
  }
  if (WANT_CV) *cv = c;
  return e;
}
// compute_mixture_ve_energy / compute_mixture_vib_energy, src/Thermo.cu:716-744
template <int M>
DEV double mix_ve_energy(const Tab *__restrict__ tb, double t, const double *y, double *cv_ve) {
  constexpr int NS = kNs(M);
  double e = 0.0, c = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double ci = 0.0;
    const double ei = cv_ve ? ve_energy1<true>(tb, i, t, &ci) : ve_energy1<false>(tb, i, t, nullptr);
    e += y[i] * ei;
    c += y[i] * ci;
  }
  if (cv_ve) *cv_ve = c;
  return e;
}
template <int M>
DEV double mix_vib_energy(const Tab *__restrict__ tb, double t, const double *y) {
  constexpr int NS = kNs(M);
  double e = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) e += y[i] * vib_energy1(tb, i, t);
  return e;
}
// compute_mixture_tr_energy, src/Thermo.cu:746-770: E_tr with the
// translational-rotational c_v and c_p and the mixture gas constant
template <int M>
DEV double mix_tr_energy(const Tab *__restrict__ tb, double t, const double *y, double &cv_tr, double &cp_tr,
                         double &r_mix) {
  constexpr int NS = kNs(M);
  double h[NS], cp[NS];
  thermo_h_cp<M>(tb, t, h, cp);
  double e = 0.0, cpm = 0.0, rm = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    cpm += y[i] * cp[i];
    rm += y[i] * tb->gas_const[i];
    e += y[i] * (h[i] - tb->gas_const[i] * t);
  }
  double cv_ve = 0.0;
  const double eve = mix_ve_energy<M>(tb, t, y, &cv_ve);
  cv_tr = (cpm - rm) - cv_ve;
  cp_tr = cpm - cv_ve;
  r_mix = rm;
  return e - eve;
}
// invert_tve_from_eve, src/Thermo.cu:772-787 (Newton, at most 80 steps)
// The loop exits through a flag rather than through breaks: the same
// iterations, with one exit edge per lane.
template <int M>
DEV double invert_tve(const Tab *__restrict__ tb, double eve_target, const double *y, double t_init) {
  double tve = fmax(t_init, 1.0);
  bool go = true;
  for (int it = 0; it < 80 && go; ++it) {
    double cv_ve = 0.0;
    const double eve = mix_ve_energy<M>(tb, tve, y, &cv_ve);
    const double res = eve - eve_target;
    if (fabs(res) < 1e-8 * fmax(1.0, fabs(eve_target))) {
      go = false;
    } else {
      const double next = fmax(1.0, tve - res / fmax(cv_ve, 1e-8));
      if (fabs(next - tve) < 1e-8 * fmax(1.0, tve)) go = false;
      tve = next;
    }
  }
  return tve;
}
// Millikan-White pair time with Park's correction, src/Thermo.cu:805-854
DEV double mw_pair_tau(const Tab *__restrict__ tb, int iv, int ip, double t, double p) {
  if (p <= 1e-16 || t <= 1.0) return 1e30;
  const double mv = 1.0 / fmax(tb->imw[iv], 1e-20), mp = 1.0 / fmax(tb->imw[ip], 1e-20);
  const double mu = mv * mp / fmax(mv + mp, 1e-20);
  double a = tb->lt_a[iv][ip], b = tb->lt_b[iv][ip];
  if (a <= 0.0 || b <= 0.0) {
    a = 1.16e-3 * sqrt(fmax(mu, 1e-20)) * pow(fmax(tb->theta_v[iv], 0.0), 4.0 / 3.0);
    b = 0.015 * pow(fmax(mu, 1e-20), 0.25);
  }
  const double ex = a * (pow(t, -1.0 / 3.0) - b) - 18.421;
  return fmax(exp(ex) * kPatm / p, 1e-30);
}
template <int M>
DEV double species_vt_tau(const Tab *__restrict__ tb, int iv, double rho, double t, double p, const double *y) {
  constexpr int NS = kNs(M);
  double conc = 0.0, den = 0.0;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const double yj = fmax(y[j], 0.0);
    if (yj <= 0.0) continue;
    const double mole = yj / (1.0 / fmax(tb->imw[j], 1e-20));
    conc += mole;
    den += mole / mw_pair_tau(tb, iv, j, t, p);
  }
  if (conc <= 0.0 || den <= 0.0) return 1e30;
  const double tau_mw = conc / den;
  const double sigma = fmax(tb->park_sigma[iv], 1e-30) * (2.5e9 / fmax(t * t, 1.0));
  const double n_total = rho * conc * 6.02214076e26;
  const double mv = 1.0 / fmax(tb->imw[iv], 1e-20);
  const double cs = sqrt(fmax(8.0 * kRu * t / (kPi * mv), 1e-30));
  const double tau_park = 1.0 / fmax(sigma * cs * n_total, 1e-30);
  return fmax(tau_mw + tau_park, 1e-30);
}
// compute_vt_relaxation_source, src/Thermo.cu:856-901 (the source itself is
// not needed: the update uses the equilibrium energy and the effective time)
template <int M>
DEV void vt_relaxation_time(const Tab *__restrict__ tb, double rho, double t, double tve, const double *y,
                            double &eve_eq, double &tau_eff) {
  constexpr int NS = kNs(M);
  double rm = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) rm += fmax(y[i], 0.0) * tb->gas_const[i];
  const double p = rho * rm * fmax(t, 1.0);
  double eq = 0.0, em = 0.0, src = 0.0;
  for (int i = 0; i < NS; ++i) {
    if (tb->theta_v[i] <= 0.0) continue;
    const double yi = fmax(y[i], 0.0);
    if (yi <= 0.0) continue;
    const double e_eq = vib_energy1(tb, i, t), e_i = vib_energy1(tb, i, tve), d = e_eq - e_i;
    eq += yi * e_eq;
    em += yi * e_i;
    if (fabs(d) <= 1e-16) continue;
    const double tau = species_vt_tau<M>(tb, i, rho, t, p, y);
    if (tau < 1e29) src += rho * yi * d / tau;
  }
  eve_eq = eq;
  const double dm = eq - em;
  tau_eff = (fabs(dm) > 1e-16 && fabs(src) > 1e-16) ? fmax(rho * dm / src, 1e-30) : 1e30;
}

// sync_two_temperature_state_from_sv, src/FieldOperation.cuh:28-57
template <int M>
DEV void sync_two_temperature(const Zone &z, const Tab *__restrict__ tb, long p) {
  constexpr int NS = kNs(M);
  const long sz = z.sz;
  double y[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = z.sv[p + l * sz];
  const double t = fmax(z.bv[p + 5 * sz], 1.0);
  double tve = z.tve[p];
  const double max_guess = fmax(1.0e7, 1000.0 * t);
  if (!isfinite(tve) || tve <= 0.0 || tve > max_guess) tve = t;
  double eve = z.sv[p + NS * sz];
  if (!isfinite(eve) || eve <= 0.0) {
    eve = mix_ve_energy<M>(tb, tve, y, nullptr);
    z.sv[p + NS * sz] = eve;
  } else {
    const double back = invert_tve<M>(tb, eve, y, tve);
    if (isfinite(back) && back > 0.0 && back <= max_guess) {
      tve = back;
    } else {
      tve = t;
      z.sv[p + NS * sz] = mix_ve_energy<M>(tb, tve, y, nullptr);
    }
  }
  z.tve[p] = tve;
}

// compute_temperature_and_pressure, two-temperature branch,
// src/FieldOperation.cu:24-64: Tve from the E_ve scalar, then T from
// e - E_ve = E_tr(T).  eve is the scalar on entry; when Tve cannot be
// recovered from it, it is reset to E_ve(T) and true is returned.
// Kept out of line: inlined into the update at -O3, this hipcc (ROCm 7.2)
// returned T = 1 K for some lanes of a wave whose lanes alone, at -O1, or
// through a call all give the converged T.
template <int M>
__device__ __noinline__ bool tt_temperature_pressure(const Zone &z, const Tab *__restrict__ tb, long p, double density, double e,
                                 const double *y, double gas_const, double &eve_io) {
  const long sz = z.sz;
  double eve = fmax(eve_io, 0.0);
  const double t_old = z.bv[p + 5 * sz];
  double tve = z.tve[p];
  if (tve <= 0) tve = fmax(t_old, 1.0);
  tve = invert_tve<M>(tb, eve, y, tve);
  const double max_tve = fmax(1.0e7, 1000.0 * fmax(t_old, 1.0));
  bool reset = false;
  if (!isfinite(eve) || !isfinite(tve) || tve <= 0.0 || tve > max_tve) {
    tve = fmax(t_old, 1.0);
    eve = mix_ve_energy<M>(tb, tve, y, nullptr);
    eve_io = eve;
    reset = true;
  }
  z.tve[p] = tve;
  double err = 1, t = fmax(t_old, 1.0);
  for (int iter = 0; iter < 1000 && err > 1e-5; ++iter) {
    double cv_tr, cp_tr, rm;
    const double e_tr = mix_tr_energy<M>(tb, t, y, cv_tr, cp_tr, rm);
    const double t1 = fmax(1.0, t - (e_tr - (e - eve)) / fmax(cv_tr, 1e-8));
    err = fabs(1 - t1 / t);
    t = t1;
  }
  z.bv[p + 5 * sz] = t;
  z.bv[p + 4 * sz] = density * t * gas_const;
  return reset;
}

// ---------------------------------------------------------------------------
// total energy from primitives, compute_cv_from_bv_1_point + compute_total_energy
// (1T), src/FieldOperation.cuh:90-161
// ---------------------------------------------------------------------------
template <int M>
DEV void cv_from_bv_point(const Zone &z, const Tab *__restrict__ tb, long p) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long sz = z.sz;
  // Every input is read before the first store: the planes never overlap, but the
  // compiler cannot know that and would wait for each store before the next load.
  const double rho = z.bv[p], u = z.bv[p + sz], v = z.bv[p + 2 * sz], w = z.bv[p + 3 * sz];
  const double pres = z.bv[p + 4 * sz], T = z.bv[p + 5 * sz];
  double y[NS > 0 ? NS : 1];
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = z.sv[p + l * sz];
  // (two temperatures: reads bv, sv and Tve, writes only E_ve and Tve)
  if constexpr (kTT(M)) sync_two_temperature<M>(z, tb, p);
  z.cv[p] = rho;
  z.cv[p + sz] = rho * u;
  z.cv[p + 2 * sz] = rho * v;
  z.cv[p + 3 * sz] = rho * w;
#pragma unroll
  for (int l = 0; l < NS; ++l) z.cv[p + (5 + l) * sz] = rho * y[l];
  const double V2 = u * u + v * v + w * w;
  double E = 0.5 * V2;
  if constexpr (kTT(M)) {
    const double eve = z.sv[p + NS * sz];
    z.cv[p + (5 + NS) * sz] = rho * eve;
    double h[NS], cp[NS];
    thermo_h_cp<M>(tb, T, h, cp);
#pragma unroll
    for (int l = 0; l < NS; ++l) E += (h[l] - tb->gas_const[l] * T) * y[l];
    E -= mix_ve_energy<M>(tb, T, y, nullptr);
    E += eve;
    E *= rho;
  } else if constexpr (NS > 0) {
    double h[NS];
    thermo_h<M>(tb, T, h);
#pragma unroll
    for (int l = 0; l < NS; ++l) E += h[l] * y[l];
    E *= rho;
    E -= pres;
  } else {
    E *= rho;
    E += pres / (kGammaAir - 1);
  }
  z.cv[p + 4 * sz] = E;
}

DEV void padded_ijk(const Zone &z, long t, int &i, int &j, int &k) {
  const long nxp = z.mx + 2 * z.ngg, nyp = z.my + 2 * z.ngg;
  i = static_cast<int>(t % nxp) - z.ngg;
  const long r = t / nxp;
  j = static_cast<int>(r % nyp) - z.ngg;
  k = static_cast<int>(r / nyp) - z.ngg;
}
DEV void interior_ijk(const Zone &z, long t, int &i, int &j, int &k) {
  i = static_cast<int>(t % z.mx);
  const long r = t / z.mx;
  j = static_cast<int>(r % z.my);
  k = static_cast<int>(r / z.my);
}

__global__ void k_fill(double *p, long n, double v) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t < n) p[t] = v;
}
__global__ void k_copy(double *__restrict__ d, const double *__restrict__ s, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t < n) d[t] = s[t];
}
__global__ void k_scaled_copy(double *__restrict__ d, const double *__restrict__ s, double a, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t < n) d[t] = a * s[t];
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_cv_from_bv(Zone z, const Tab *__restrict__ tb, long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  padded_ijk(z, t, i, j, k);
  cv_from_bv_point<M>(z, tb, z.ix(i, j, k));
}

// ---------------------------------------------------------------------------
// transport: Sutherland (Air) and mixture-averaged mu, kappa, rho*D
// (src/Transport.cu:6-150, Omega_D fit 182-184)
// ---------------------------------------------------------------------------
// 1/x from v_rcp_f64 and two Newton steps (each squares the relative error),
// within an ulp of the correctly rounded reciprocal for the normal, finite,
// nonzero arguments it is used on (WENO weight sums, transport sums); 5
// instructions instead of the IEEE division sequence.
DEV double fast_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) from v_rsq_f64 and two Newton steps y (3/2 - x y^2 / 2), within an
// ulp or two for the positive normal arguments (viscosities) it is used on.
DEV double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double e = fma(-hx * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-hx * y, y, 0.5);
  return fma(y, e, y);
}

DEV double sutherland(double t) { return 1.7894e-5 * pow(t / 288.16, 1.5) * (288.16 + 110) / (t + 110); }

// Collision-integral fit for diffusion, src/Transport.cu:182-184:
// Omega_D(t*) = 1.0548 t*^-0.15504 + (t* + 0.55909)^-2.1705 with t* = T c.
// The first power splits into T^-0.15504 (once per point) times the pair
// constant 1.0548 c^-0.15504; the second takes the short log / exp of
// fastmath.h (within a few ulp), evaluated for every species pair.
DEV double omega_d_fit(double t_red, double t_pw, double pair_pw) {
  return t_pw * pair_pw + fm_exp(-2.1705 * fm_log(t_red + 0.55909));
}

// Mixture-averaged mu, kappa, rho*D (src/Transport.cu:43-150).  Powers of the
// reduced temperature split into a power of T per point and per-species or
// per-pair constants computed once with the tables.
template <int M>
DEV void transport(const Tab *__restrict__ tb, double temperature, double mw_total, const double *y,
                   const double *cp, double &mul, double &kappa, double *rhoD) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  double X[NS], vis[NS], rsv[NS];
  const double sqrt_t = sqrt(temperature);
  const double ln_t = log(temperature);
  const double t_pw145 = exp(-0.145 * ln_t), t_pw15504 = exp(-0.15504 * ln_t);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    X[l] = y[l] * mw_total * tb->imw[l];
    const double t_dl = temperature * tb->LJ_potent_inv[l];
    const double a = t_dl + 0.5;
    vis[l] = tb->vis_coeff[l] * sqrt_t * fast_rcp(1.147 * (t_pw145 * tb->pw_lj[l]) + fast_rcp(a * a));
    rsv[l] = fast_rsqrt(vis[l]);
  }
  const double t_3over2 = temperature * sqrt_t;
  const double inv_t32 = 1.0 / t_3over2;
  double viscosity = 0, conductivity = 0;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    double R = tb->gas_const[m];
    double lambda = 15 * 0.25 * vis[m] * R;
    const int geom = tb->geometry[m];
    if (geom == 1 || geom == 2) {
      const double t_red = temperature * tb->kb_over_eps_jk[m][m];
      const double tRedInv = fast_rcp(t_red);
      const double tmp = sqrt(tRedInv * kPi) * kPi;
      const double FT = 1 + 0.5 * tmp + (2 + 0.25 * kPi * kPi) * tRedInv + tmp * tRedInv;
      const double rD = tb->binary_diffusivity_coeff[m][m] * t_3over2 *
                        fast_rcp(omega_d_fit(t_red, t_pw15504, tb->pw_kbe[m][m]) * R * temperature);
      const double zft = kPi * tb->ZRotF298[m] * fast_rcp(FT);
      if (geom == 1) {
        const double ADivPiB = (2.5 * vis[m] - rD) * fast_rcp(zft * vis[m] + 10.0 / 3 * vis[m] + 2.0 * rD);
        lambda += -5 * ADivPiB * vis[m] * R + rD * cp[m] + rD * (2 * ADivPiB - 2.5) * R;
      } else {
        const double ADivPiB = (2.5 * vis[m] - rD) * fast_rcp(zft * vis[m] + 5 * vis[m] + 2.0 * rD);
        lambda += -7.5 * ADivPiB * vis[m] * R + rD * cp[m] + rD * (3 * ADivPiB - 2.5) * R;
      }
    }
    double sum = 0;
    R = vis[m] * rsv[m];  // sqrt(mu_m)
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double pf = 1.0;
      if (m != k) {
        const double numer = 1.0 + rsv[k] * (R * tb->WjDivWi_to_One4th[m][k]);
        pf = numer * numer * tb->sqrt_WiDivWjPl1Mul8[m][k];
      }
      sum += pf * X[k];
    }
    viscosity += vis[m] * X[m] * fast_rcp(sum);
    conductivity += lambda * X[m] * fast_rcp(1.065 * sum - 0.065 * X[m]);
  }
  mul = viscosity;
  kappa = conductivity;

  constexpr double eps = 1e-12;
  double num[NS], den[NS];
  {
    // The numerator (sum of all (X+eps)W minus the species' own term) cancels
    // catastrophically where one species is nearly pure; without contraction
    // these are the reference's operations exactly.
#pragma clang fp contract(off)
    double s = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double tmp = (X[l] + eps) * (1.0 / tb->imw[l]);
      num[l] = tmp;
      s += tmp;
      den[l] = 0;
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) num[l] = s - num[l];
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) {
#pragma unroll
    for (int k = l + 1; k < NS; ++k) {
      const double t_red = temperature * tb->kb_over_eps_jk[l][k];
      const double tmp = omega_d_fit(t_red, t_pw15504, tb->pw_kbe[l][k]) * tb->inv_bdc[l][k] * inv_t32;
      den[l] += (X[k] + eps) * tmp;
      den[k] += (X[l] + eps) * tmp;
    }
  }
  const double inv = 1.0 / (temperature * kRu);
#pragma unroll
  for (int l = 0; l < NS; ++l) rhoD[l] = num[l] * inv * fast_rcp(den[l]);
}

// update_physical_properties (1T) over the padded extent, src/FieldOperation.cuh:183-247
template <int M>
DEV void properties_point(const Zone &z, const Tab *__restrict__ tb, int transport_ghosts, long t) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  int i, j, k;
  padded_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz;
  const double T = z.bv[p + 5 * sz];
  const double V = norm3(z.bv[p + sz], z.bv[p + 2 * sz], z.bv[p + 3 * sz]);
  if constexpr (NS > 0) {
    double cp[NS], y[NS];
    thermo_cp<M>(tb, T, cp);
    double imw = 0, cpt = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = z.sv[p + l * sz];
      imw = fma(y[l], tb->imw[l], imw);
      cpt = fma(y[l], cp[l], cpt);
    }
    const double mw = 1 / fmax(imw, 1e-12);
    const double R = kRu * imw;
    if constexpr (kTT(M)) {
      // translational-rotational gamma and sound speed, src/FieldOperation.cuh:207-229
      double cv_tr, cpm, rm;
      mix_tr_energy<M>(tb, T, y, cv_tr, cpm, rm);
      const double gam = cpm / fmax(cv_tr, 1e-8);
      z.gam[p] = gam;
      const double c = sqrt(fmax(gam * rm * T, 1e-12));
      z.acoustic[p] = c;
      z.mach[p] = V / c;
      if (z.tve[p] <= 0) z.tve[p] = T;
    } else {
      const double gam = cpt / (cpt - R);
      z.gam[p] = gam;
      const double c = sqrt(gam * R * T);
      z.acoustic[p] = c;
      z.mach[p] = V / c;
    }
    // The 8th-order viscous fluxes and the time-step bound read mu, kappa
    // and rho*D at interior nodes only; the 2nd-order face scheme also reads
    // them one node outside, so there the padded extent is covered.
    if (transport_ghosts || (i >= 0 && i < z.mx && j >= 0 && j < z.my && k >= 0 && k < z.mz)) {
      double rd[NS], mul, kap;
      transport<M>(tb, T, mw, y, cp, mul, kap, rd);
      z.mul[p] = mul;
      z.kappa[p] = kap;
      if constexpr (kTT(M)) {
        double cv_ve = 0.0;
        mix_ve_energy<M>(tb, z.tve[p], y, &cv_ve);
        z.kappa_ve[p] = kap * cv_ve / fmax(cpt, 1e-8);
      }
#pragma unroll
      for (int l = 0; l < NS; ++l) z.rhoD[p + l * sz] = rd[l];
    }
  } else {
    constexpr double c_temp = kGammaAir * kRu / kMwAir;
    z.mul[p] = sutherland(T);
    z.mach[p] = V / sqrt(c_temp * T);
  }
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_properties(Zone z, const Tab *__restrict__ tb, int transport_ghosts,
                                                       long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  properties_point<M>(z, tb, transport_ghosts, t);
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_properties_multi(const Zone *__restrict__ zs, const Tab *__restrict__ tb,
                                                             int transport_ghosts) {
  const Zone &z = zs[blockIdx.y];
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  const long n = static_cast<long>(z.mx + 2 * z.ngg) * (z.my + 2 * z.ngg) * (z.mz + 2 * z.ngg);
  if (t >= n) return;
  properties_point<M>(z, tb, transport_ghosts, t);
}

// ---------------------------------------------------------------------------
// SSP-RK3 update + primitive recovery, src/RK.cuh:277-367 with
// compute_temperature_and_pressure (1T), src/FieldOperation.cu:4-94
// ---------------------------------------------------------------------------
// dtp: the step's dt in device memory (a captured step reads it there), else dt;
// dtc: a dt per interior point (batched shock tubes, each on its own step).
DEV bool pos_finite(double x) { return x > 0.0 && x < INFINITY; }

// The states at which the reference's residual check stops a run
// (src/Residual.cu:97-118, 211-243): a non-finite or non-positive density,
// pressure or temperature, non-finite velocity or species.  Every writer
// stores the same value, so a plain store is enough.
DEV void flag_state(const Zone &z, bool ok) {
  if (!ok && z.bad) *z.bad = 1;
}

// WU: Wu's splitting's RK sub-step, whose increment is dq - Cn
// (update_cv_and_bv_rk_wu, src/StrangSplitting.cu:508-566).  save_qn (the
// first stage of a recorded step): cv still holds the step's start state, so
// it serves as qn and is saved there, in place of a separate copy of cv.
template <int M, bool WU = false, bool SAVE_QN = false>
DEV void update_point(const Zone &z, const Tab *__restrict__ tb, double dt, int rk, long t) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), q = t, sz = z.sz, dsz = z.dsz;
  constexpr int NV = kNv(M);
  const double a = c_rk_a[rk], b = c_rk_b[rk], c = c_rk_c[rk];
  const double dt_div_jac = dt / z.jac[p];
  double cv[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double inc = WU ? z.dq[q + l * dsz] - z.cn[q + l * dsz] : z.dq[q + l * dsz];
    const double cvl = z.cv[p + l * sz];
    double qnl;
    if constexpr (SAVE_QN) {
      qnl = cvl;
      z.qn[p + l * sz] = cvl;
    } else {
      qnl = z.qn[p + l * sz];
    }
    cv[l] = a * qnl + b * cvl + c * dt_div_jac * inc;
  }
  if (z.mz == 1) cv[3] = 0;
  const double density_inv = 1.0 / cv[0];
  const double u = cv[1] * density_inv, v = cv[2] * density_inv, w = cv[3] * density_inv;
  z.bv[p] = cv[0];
  z.bv[p + sz] = u;
  z.bv[p + 2 * sz] = v;
  z.bv[p + 3 * sz] = w;
  if constexpr (NS > 0) {
    double sum = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      if (cv[5 + l] < 0)
        cv[5 + l] = 0;
      else
        sum += cv[5 + l];
    }
    if constexpr (kTT(M)) cv[5 + NS] = fmax(cv[5 + NS], 0.0);
    sum = 1.0 / sum;
    const double ratio = cv[0] * sum;
    double y[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      y[l] = cv[5 + l] * sum;
      cv[5 + l] *= ratio;
      z.sv[p + l * sz] = y[l];
    }
    double gas_const = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) gas_const += y[l] * tb->gas_const[l];
    const double e = cv[4] / cv[0] - 0.5 * (u * u + v * v + w * w);
    if constexpr (kTT(M)) {
      double eve = cv[5 + NS] * density_inv;
      z.sv[p + NS * sz] = eve;
      if (tt_temperature_pressure<M>(z, tb, p, cv[0], e, y, gas_const, eve)) {
        z.sv[p + NS * sz] = eve;
        cv[5 + NS] = cv[0] * eve;
      }
#pragma unroll
      for (int l = 0; l < NV; ++l) z.cv[p + l * sz] = cv[l];
      flag_state(z, pos_finite(cv[0]) && pos_finite(sum) && isfinite(u) && isfinite(v) && isfinite(w) &&
                        pos_finite(z.bv[p + 4 * sz]) && pos_finite(z.bv[p + 5 * sz]) && isfinite(eve));
      return;
    }
    // Newton on e(T) warm-started from the stored T, relative tolerance 1e-5
    double err = 1, tt = z.bv[p + 5 * sz];
    int iter = 0;
    while (err > 1e-5 && iter++ < 1000) {
      double h[NS], cp[NS];
      thermo_h_cp<M>(tb, tt, h, cp);
      double cp_tot = 0, hh = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        cp_tot += cp[l] * y[l];
        hh += h[l] * y[l];
      }
      const double e_t = hh - gas_const * tt;
      const double cvv = cp_tot - gas_const;
      const double t1 = tt - (e_t - e) / cvv;
      err = fabs(1 - t1 / tt);
      tt = t1;
    }
    const double pr = cv[0] * tt * gas_const;
    z.bv[p + 5 * sz] = tt;
    z.bv[p + 4 * sz] = pr;
    flag_state(z, pos_finite(cv[0]) && pos_finite(sum) && isfinite(u) && isfinite(v) && isfinite(w) &&
                      pos_finite(pr) && pos_finite(tt));
  } else {
    const double V2 = u * u + v * v + w * w;
    const double pr = (kGammaAir - 1) * (cv[4] - 0.5 * cv[0] * V2);
    const double tt = pr * kMwAir * density_inv / kRu;
    z.bv[p + 4 * sz] = pr;
    z.bv[p + 5 * sz] = tt;
    flag_state(z, pos_finite(cv[0]) && isfinite(u) && isfinite(v) && isfinite(w) && pos_finite(pr) &&
                      pos_finite(tt));
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) z.cv[p + l * sz] = cv[l];
}

template <int M, bool WU = false, bool SAVE_QN = false>
__global__ __launch_bounds__(kBlock) void k_update_rk(Zone z, const Tab *__restrict__ tb, double dt,
                                                      const double *__restrict__ dtp,
                                                      const double *__restrict__ dtc, int rk, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  if (dtc)
    dt = dtc[t];
  else if (dtp)
    dt = *dtp;
  update_point<M, WU, SAVE_QN>(z, tb, dt, rk, t);
}

// The update of every block of a recorded multi-block step in one launch
// (blockIdx.y = block; the zones in device memory, dt in *dtp).
template <int M, bool SAVE_QN>
__global__ __launch_bounds__(kBlock) void k_update_rk_multi(const Zone *__restrict__ zs, const Tab *__restrict__ tb,
                                                            const double *__restrict__ dtp, int rk) {
  const Zone &z = zs[blockIdx.y];
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<long>(z.mx) * z.my * z.mz) return;
  update_point<M, false, SAVE_QN>(z, tb, *dtp, rk, t);
}

// ---------------------------------------------------------------------------
// V-T relaxation after the update, src/RK.cuh:20-28 and 200-203,
// src/FieldOperation.cuh:60-88: the vibrational energy relaxes exactly
// (exponentially) toward its value at T over the stage's dt, then Tve and T
// are recovered again.  Interior points.
// ---------------------------------------------------------------------------
template <int M>
__global__ __launch_bounds__(kBlock) void k_vt_relax(Zone z, const Tab *__restrict__ tb, double dt_stage,
                                                     const double *__restrict__ dtp, double scale, long n) {
  constexpr int NS = kNs(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  if (dtp) dt_stage = scale * *dtp;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz;
  const double density = z.cv[p];
  if (density <= 0.0) return;
  const double T = z.bv[p + 5 * sz];
  const double tve = z.tve[p] > 0.0 ? z.tve[p] : fmax(T, 1.0);
  double y[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = z.sv[p + l * sz];
  double ev_eq, tau;
  vt_relaxation_time<M>(tb, density, T, tve, y, ev_eq, tau);
  if (tau >= 1e29 || !isfinite(tau)) return;
  const double rho_eve_old = fmax(z.cv[p + (5 + NS) * sz], 0.0);
  const double alpha = exp(-dt_stage / fmax(tau, 1e-30));
  const double ev_old = mix_vib_energy<M>(tb, tve, y);
  const double ev_new = ev_old * alpha + ev_eq * (1.0 - alpha);
  const double rho_eve = fmax(rho_eve_old + density * (ev_new - ev_old), 0.0);
  z.cv[p + (5 + NS) * sz] = rho_eve;
  double eve = rho_eve / density;
  z.sv[p + NS * sz] = eve;
  z.tve[p] = invert_tve<M>(tb, eve, y, tve);
  double gas_const = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) gas_const += y[l] * tb->gas_const[l];
  const double rho = z.bv[p], u = z.bv[p + sz], v = z.bv[p + 2 * sz], w = z.bv[p + 3 * sz];
  const double e = z.cv[p + 4 * sz] / rho - 0.5 * (u * u + v * v + w * w);
  if (tt_temperature_pressure<M>(z, tb, p, rho, e, y, gas_const, eve)) {
    z.sv[p + NS * sz] = eve;
    z.cv[p + (5 + NS) * sz] = rho * eve;
  }
}

// ---------------------------------------------------------------------------
// finite-rate chemistry, explicit, src/FiniteRateChem.cu:1406-1681 (generic
// mechanism path).  Reactions are processed one at a time so per-reaction
// values stay scalar; every species sum keeps the reference's reaction order.
// Stoichiometry is read from per-reaction sparse lists in species order (the
// dense sums minus their zero terms); the species index is the same for the
// whole wave, so per-species values stay in registers and are addressed
// indirectly.  Integer stoichiometric powers are evaluated as products (exact
// for 1 and 2).
// ---------------------------------------------------------------------------
// reaction_temperature with smooth_reaction_temperature, src/FiniteRateChem.cu:1490-1503
DEV double reaction_temperature(double t, double tve, double a, double b) {
  const double trx = pow(fmax(t, 1.0), a) * pow(fmax(tve, 1.0), b);
  return 0.5 * (trx + 800.0 + sqrt((trx - 800.0) * (trx - 800.0) + 80.0 * 80.0));
}
// G/RT of one species (same arithmetic as thermo_gibbs)
template <int M>
DEV double gibbs1(const Tab *__restrict__ tb, int i, double t) {
  if constexpr (kNasa9(M)) {
    const int n = tb->n_range[i];
    const double lo = tb->t_cut[i][0], hi = tb->t_cut[i][n];
    if (t < lo || t > hi) return nasa9_g_at(tb->poly9[i][t < lo ? 0 : n - 1], t < lo ? lo : hi);
    return nasa9_g_at(tb->poly9[i][nasa9_range(tb, i, t)], t);
  } else {
    const double tl = tb->t_low[i], th = tb->t_high[i];
    const bool out = t < tl || t > th;
    const double tt = out ? (t < tl ? tl : th) : t;
    const double *c = out ? (t < tl ? tb->low_coeff[i] : tb->high_coeff[i])
                          : (t < tb->t_mid[i] ? tb->low_coeff[i] : tb->high_coeff[i]);
    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
    return c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
           c[5] * tt_inv - c[6];
  }
}

DEV double ipow(double c, int e) {
  if (e == 0) return 1.0;
  if (e == 1) return c;
  if (e == 2) return c * c;
  return pow(c, static_cast<double>(e));
}

// Arrhenius A T^b exp(-Ea/T) as one exponential of b log T - Ea/T.
DEV double arrhenius_ln(double ln_t, double inv_t, double A, double b, double Ea) {
  return A * exp(b * ln_t - Ea * inv_t);
}

// Forward and backward rates of progress q1, q2 of reaction r at T (and Tve
// when trx) from the concentrations c (mol/cm^3) and G/RT g of the species
// (src/FiniteRateChem.cu:1490-1663): the per-reaction part of chem_rates,
// shared by the per-point kernels and the reactor kernel that spreads the
// reactions of one reactor over a wave's lanes.
// log of p_atm / (R_u T) in mol/cm^3: the standard-state concentration of Kc
DEV double ln_std_conc(double ln_t) {
  constexpr double temp_p = kPatm / kRu * 1e-3;
  return log(temp_p) - ln_t;
}

template <int M>
DEV void reaction_progress(const Tab *__restrict__ tb, int r, double T, double tve, bool trx, double ln_t,
                           double inv_t, double ln_temp_t, const double *c, const double *g, double &q1,
                           double &q2) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  constexpr double temp_p = kPatm / kRu * 1e-3;
  constexpr double kLn10 = 2.302585092994046, kInvLn10 = 0.4342944819032518;
  const Reac &R = tb->reac[r];
  const int type = tb->reac_type[r];
  double thf = T, lt = ln_t, it = inv_t;
  if (kTT(M) && trx) {
    thf = reaction_temperature(T, tve, tb->tcf_a[r], tb->tcf_b[r]);
    lt = log(thf);
    it = 1.0 / thf;
  }
  double kf = arrhenius_ln(lt, it, tb->A[r], tb->b[r], tb->Ea[r]);
  if (type == 3) {
    kf += arrhenius_ln(lt, it, tb->A2[r], tb->b2[r], tb->Ea2[r]);
  } else if (type > 3) {
    double cc = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) cc += c[l] * tb->third_body[r][l];
    if (type == 4) {
      kf *= cc;
    } else {
      const double kf_low = arrhenius_ln(lt, it, tb->A2[r], tb->b2[r], tb->Ea2[r]);
      const double kf_high = kf;
      if (kf_high < 1e-25 && kf_low < 1e-25) {
        kf = 0;
      } else {
        // Lindemann / Troe fall-off (src/FiniteRateChem.cu:1530-1551)
        const double reduced_pressure = kf_low * cc / kf_high;
        double F = 1.0;
        if (type > 5) {
          double f_cent = (1 - tb->troe_alpha[r]) * exp(-thf / tb->troe_t3[r]) +
                          tb->troe_alpha[r] * exp(-thf / tb->troe_t1[r]);
          if (type == 7) f_cent += exp(-tb->troe_t2[r] * it);
          const double logFc = log(f_cent) * kInvLn10;
          const double cT = -0.4 - 0.67 * logFc;
          const double nn = 0.75 - 1.27 * logFc;
          const double logPr = log(reduced_pressure) * kInvLn10;
          const double tempo = (logPr + cT) / (nn - 0.14 * (logPr + cT));
          F = exp(kLn10 * (logFc / (1.0 + tempo * tempo)));
        }
        kf = kf_high * reduced_pressure / (1.0 + reduced_pressure) * F;
      }
    }
  }
  // backward rate (1T branch): REV sets kb, then every type other than 0/2
  // is overwritten from the equilibrium constant, as the reference does
  // (src/FiniteRateChem.cu:1591-1636)
  double kb = 0;
  if (kTT(M) && trx && type != 0) {
    // per-reaction backward temperature; REV keeps its rate, otherwise
    // kf / Kc with G/RT at that temperature (src/FiniteRateChem.cu:1560-1589)
    const double thb = reaction_temperature(T, tve, tb->tcb_a[r], tb->tcb_b[r]);
    const double lb = log(thb);
    if (tb->rev_type[r] == 1) {
      kb = arrhenius_ln(lb, 1.0 / thb, tb->A2[r], tb->b2[r], tb->Ea2[r]);
      if (type == 4) {
        double cc = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) cc += c[l] * tb->third_body[r][l];
        kb *= cc;
      }
    } else {
      double d_gibbs = 0;
      for (int m = 0; m < R.nnu; ++m) d_gibbs += gibbs1<M>(tb, R.nu_idx[m], thb) * R.nu[m];
      kb = kf * exp(d_gibbs - static_cast<double>(tb->order[r]) * (log(temp_p) - lb));
    }
  } else if (type != 0) {
    if (tb->rev_type[r] == 1) {
      kb = arrhenius_ln(ln_t, inv_t, tb->A2[r], tb->b2[r], tb->Ea2[r]);
      if (type == 4) {
        double cc = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) cc += c[l] * tb->third_body[r][l];
        kb *= cc;
      }
    }
    if (type != 2) {
      // (the species with nu_b - nu_f != 0 in index order, as the reaction's nu list holds them; a loop over
      // every species with a test the wave takes together keeps c and g in registers without indexing them)
      const int *sf = tb->stoi_f[r], *sb = tb->stoi_b[r];
      double d_gibbs = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l)
        if (sb[l] != sf[l]) d_gibbs += g[l] * static_cast<double>(sb[l] - sf[l]);
      // kb = kf / Kc with Kc = (p_atm / R_u T)^order exp(-dG/RT)
      kb = kf * exp(d_gibbs - static_cast<double>(tb->order[r]) * ln_temp_t);
    }
  }
  // rate of progress, src/FiniteRateChem.cu:1638-1663 (reactants and products in index order)
  const int *sf = tb->stoi_f[r], *sb = tb->stoi_b[r];
  q1 = 1.0;
  q2 = 1.0;
#pragma unroll
  for (int l = 0; l < NS; ++l)
    if (sf[l] != 0) q1 *= ipow(c[l], sf[l]);
  q1 *= kf;
  if (type != 0) {
#pragma unroll
    for (int l = 0; l < NS; ++l)
      if (sb[l] != 0) q2 *= ipow(c[l], sb[l]);
    q2 *= kb;
  } else {
    q2 = 0.0;
  }
}

// Creation and destruction rates (mol / (cm^3 s), summed with the
// stoichiometry) of every species at (density, T, Tve, y): forward rates,
// fall-off, backward rates from REV or Kc, rates of progress, chemical source
// (src/FiniteRateChem.cu:1406-1681).  With JAC the sums
// zz[m][n] = sum_r (nu_b - nu_f)_rm (nu_f,rn q1_r - nu_b,rn q2_r) of the source
// Jacobian (src/ConstVolumeReactor.cpp:1739-1758) accumulate too.
// (zz element (m, n) at zz[(m NS + n) ZST]: ZST > 1 keeps a wave's matrices interleaved in LDS)
template <int M, bool JAC, int ZST = 1>
DEV void chem_rates(const Tab *__restrict__ tb, double density, double T, double tve_in, const double *y,
                    double *creation, double *omega_d, double *zz) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  double c[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) c[l] = density * y[l] * tb->imw[l] * 1e-3;
  double g[NS];
  thermo_gibbs<M>(tb, T, g);
  const double ln_t = log(T), inv_t = 1.0 / T;
  const double ln_temp_t = ln_std_conc(ln_t);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    creation[l] = 0;
    omega_d[l] = 0;
  }
  if constexpr (JAC) {
#pragma unroll
    for (int l = 0; l < NS * NS; ++l) zz[l * ZST] = 0;
  }
  // two temperatures: the rates may take a per-reaction control temperature
  // T^a Tve^b (reaction_temperature, src/FiniteRateChem.cu:1490-1503)
  [[maybe_unused]] double tve = T;
  [[maybe_unused]] bool trx = false;
  if constexpr (kTT(M)) {
    tve = fmax(tve_in, 1.0);
    trx = tb->trx_model != 0;
  }
  const int nr = tb->n_reac;
  for (int r = 0; r < nr; ++r) {
    double q1, q2;
    reaction_progress<M>(tb, r, T, tve, kTT(M) && trx, ln_t, inv_t, ln_temp_t, c, g, q1, q2);
    // chemical_source accumulation, src/FiniteRateChem.cu:1665-1681, over the species taking part (either
    // coefficient not 0), looped over every species with tests the wave takes together
    const int *sfi = tb->stoi_f[r], *sbi = tb->stoi_b[r];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      if (sfi[l] == 0 && sbi[l] == 0) continue;
      const double sf = sfi[l], sb = sbi[l];
      creation[l] += q2 * sf + q1 * sb;
      omega_d[l] += q1 * sf + q2 * sb;
      if constexpr (JAC) {
#pragma unroll
        for (int l2 = 0; l2 < NS; ++l2) {
          if (sfi[l2] == 0 && sbi[l2] == 0) continue;
          zz[(l * NS + l2) * ZST] += (sb - sf) * (static_cast<double>(sfi[l2]) * q1 - static_cast<double>(sbi[l2]) * q2);
        }
      }
    }
  }
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_chemistry(Zone z, const Tab *__restrict__ tb, long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz, dsz = z.dsz;
  const double density = z.bv[p];
  const double T = z.bv[p + 5 * sz];
  double y[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) y[l] = z.sv[p + l * sz];
  double creation[NS], omega_d[NS];
  chem_rates<M, false>(tb, density, T, kTT(M) ? z.tve[p] : T, y, creation, omega_d, nullptr);
  const double jac = z.jac[p];
  double time_scale = 1e+6;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double conv = 1e+3 / tb->imw[l];
    const double cr = creation[l] * conv;
    const double od = omega_d[l] * conv;
    const double om = cr - od;
    z.dq[t + (l + 5) * dsz] += jac * om;
    if (y[l] > 1e-25 && fabs(od) > 1e-25) time_scale = fmin(time_scale, fabs(density * y[l] / od));
  }
  if constexpr (kTT(M)) {
    // the E_ve carried by the produced species, src/FiniteRateChem.cu:19-31
    double tv = z.tve[p];
    if (tv <= 0.0) tv = fmax(T, 1.0);
    double src = 0.0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double conv = 1e+3 / tb->imw[l];
      src += (creation[l] * conv - omega_d[l] * conv) * ve_energy1<false>(tb, l, tv, nullptr);
    }
    z.dq[t + (5 + NS) * dsz] += jac * src;
  }
  z.rts[p] = time_scale;
}

// ---------------------------------------------------------------------------
// 0-D constant-volume reactors for the request server: one thread per slot
// (src/ConstVolumeReactor.cpp:320-775).  Each launch advances every active
// slot by at most max_steps attempted steps; a slot whose time reaches t_end
// is marked done.  Schemes: 0 point-implicit with the diagonal source
// Jacobian (implicit_method 2), 1 ROK4E with the full Jacobian (fixed or
// adaptive steps), 3 SSP-RK3.
// ---------------------------------------------------------------------------
// update_t, src/ConstVolumeReactor.cpp:1456-1482: Newton on E = sum Y h - R T
template <int M>
DEV double reactor_temperature(const Tab *__restrict__ tb, double t0, double imw, const double *yk, double e) {
  constexpr int NS = kNs(M);
  double t = t0, err = 1e+6;
  const double r = kRu * imw;
  int iter = 0;
  while (err > 1e-8 && iter++ < 100) {
    double h[NS], cp[NS];
    thermo_h_cp<M>(tb, t, h, cp);
    double cp_tot = 0, hm = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp_tot += cp[l] * yk[l];
      hm += h[l] * yk[l];
    }
    const double t1 = t - (hm - r * t - e) / (cp_tot - r);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

// omega (kg / (m^3 s)) from the creation / destruction sums
template <int M, int ZST = 1>
DEV void reactor_source(const Tab *__restrict__ tb, double rho, double t, const double *yk, double *omega,
                        double *zz) {
  constexpr int NS = kNs(M);
  double cr[NS], od[NS];
  if (zz)
    chem_rates<M, true, ZST>(tb, rho, t, t, yk, cr, od, zz);
  else
    chem_rates<M, false>(tb, rho, t, t, yk, cr, od, nullptr);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double conv = 1e+3 / tb->imw[l];
    omega[l] = cr[l] * conv - od[l] * conv;
  }
}

// lu_decomp / lu_to_solution (src/ConstVolumeReactor.cpp:1802-1856) with
// every index compile-time: the pivot row is exchanged through selects.
// Element (m, n) of the matrix is a[(m NS + n) ST]: ST > 1 interleaves the
// matrices of a wave's lanes in LDS.
template <int NS, int ST = 1>
DEV void lu_factor(double *a, int *ipiv) {
#pragma unroll
  for (int n = 0; n < NS; ++n) {
#pragma unroll
    for (int m = n; m < NS; ++m)
#pragma unroll
      for (int t = 0; t < n; ++t) a[(m * NS + n) * ST] -= a[(m * NS + t) * ST] * a[(t * NS + n) * ST];
    int ik = n;
    double best = fabs(a[(n * NS + n) * ST]);
#pragma unroll
    for (int m = n + 1; m < NS; ++m) {
      const double v = fabs(a[(m * NS + n) * ST]);
      if (v > best) {
        best = v;
        ik = m;
      }
    }
    ipiv[n] = ik;
#pragma unroll
    for (int m = n + 1; m < NS; ++m) {
      const bool sw = ik == m;
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        const double x = a[(n * NS + t) * ST], w = a[(m * NS + t) * ST];
        a[(n * NS + t) * ST] = sw ? w : x;
        a[(m * NS + t) * ST] = sw ? x : w;
      }
    }
#pragma unroll
    for (int q = n + 1; q < NS; ++q)
#pragma unroll
      for (int t = 0; t < n; ++t) a[(n * NS + q) * ST] -= a[(n * NS + t) * ST] * a[(t * NS + q) * ST];
#pragma unroll
    for (int m = n + 1; m < NS; ++m) a[(m * NS + n) * ST] /= a[(n * NS + n) * ST];
  }
}

template <int NS, int ST = 1>
DEV void lu_apply(const double *a, const int *ipiv, double *b) {
#pragma unroll
  for (int m = 0; m < NS; ++m) {
#pragma unroll
    for (int q = m + 1; q < NS; ++q) {
      const bool sw = ipiv[m] == q;
      const double x = b[m], w = b[q];
      b[m] = sw ? w : x;
      b[q] = sw ? x : w;
    }
  }
#pragma unroll
  for (int m = 1; m < NS; ++m)
#pragma unroll
    for (int t = 0; t < m; ++t) b[m] -= a[(m * NS + t) * ST] * b[t];
  b[NS - 1] /= a[(NS * NS - 1) * ST];
#pragma unroll
  for (int m = NS - 2; m >= 0; --m) {
#pragma unroll
    for (int t = m + 1; t < NS; ++t) b[m] -= a[(m * NS + t) * ST] * b[t];
    b[m] /= a[(m * NS + m) * ST];
  }
}

// threads per block of the reactor kernel (one slot per thread)
constexpr int kReactorBlock = 64;

// Why a reactor slot cannot go on (0: it can, or it has reached t_end): a
// non-finite temperature, time, step or partial density; a step that is not
// positive or too small to move the time; or the request's step cap reached.
template <int NS>
DEV int reactor_fault(double t, double time, double h, double t_end, const double *rhoY, long tries, long cap) {
  bool finite = isfinite(t) && isfinite(time) && isfinite(h);
#pragma unroll
  for (int l = 0; l < NS; ++l) finite = finite && isfinite(rhoY[l]);
  if (!finite) return COREFL_FAIL_NONFINITE;
  if (time >= t_end) return 0;
  const double hs = fmin(h, t_end - time);
  if (!(hs > 0) || time + hs == time) return COREFL_FAIL_STEP;
  if (cap > 0 && tries >= cap) return COREFL_FAIL_CAP;
  return 0;
}

// ROK4E tableau (the method's published coefficients, as src/Define.h:33-48)
constexpr double kRkGamma = 0.572816062482135;
constexpr double kRkA21 = 0.432364435748567, kRkA31 = -0.51421131687617, kRkA32 = 1.38227114461736;
constexpr double kRkG21 = -0.602765307997356 / kRkGamma, kRkG31 = -1.389195789724843 / kRkGamma,
                 kRkG32 = 1.072950969011413 / kRkGamma, kRkG41 = 0.992356412977094 / kRkGamma,
                 kRkG42 = -1.390032613873701 / kRkGamma, kRkG43 = -0.440875890223325 / kRkGamma;
constexpr double kRkB1 = 0.194335256262729, kRkB2 = 0.483167813989227, kRkB4 = 0.322496929748044;
constexpr double kRkE1 = -0.217819895945721 - kRkB1, kRkE2 = 1.03130847478467 - kRkB2, kRkE3 = 0.186511421161047,
                 kRkE4 = -kRkB4;

template <int M>
__global__ __launch_bounds__(kReactorBlock) void k_reactors(ReactorPool P, const Tab *__restrict__ tb,
                                                            int max_steps) {
  constexpr int NS = kNs(M);
  // ROK4E's iteration matrix in LDS (a scratch array otherwise), the lanes' matrices interleaved
  __shared__ double lu_lds[NS * NS * kReactorBlock];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.cap || P.status[s] != 1) return;
  const long cap = P.cap;
  double rhoY[NS], yk[NS], omega[NS];
  const double rho = P.rho[s], e = P.e[s], t_end = P.t_end[s], rtol = P.rtol[s], atol = P.atol[s];
  const int scheme = P.scheme[s], adaptive = P.adaptive[s], conserve = P.conserve[s];
  double t = P.t[s], time = P.time[s], h = P.h[s], err_last = P.err_last[s];
  long steps = P.steps[s], rejected = P.rejected[s];
  const long step_cap = P.max_steps[s];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    rhoY[l] = P.rhoY[l * cap + s];
    yk[l] = rhoY[l] / rho;
  }
  int fail = 0;
  for (int it = 0; it < max_steps; ++it) {
    fail = reactor_fault<NS>(t, time, h, t_end, rhoY, steps + rejected, step_cap);
    if (fail || time >= t_end) break;
    bool accepted = true;
    if (scheme == 1) {
      double k1[NS], k2[NS], k3[NS], k4[NS], ys[NS];
      // this lane's matrix (element stride kReactorBlock): the source Jacobian, then in place the
      // iteration matrix
      double *a = lu_lds + threadIdx.x;
      int ipiv[NS];
      reactor_source<M, kReactorBlock>(tb, rho, t, yk, omega, a);
      const double hs = fmin(h, t_end - time);
#pragma unroll
      for (int m = 0; m < NS; ++m)
#pragma unroll
        for (int n = 0; n < NS; ++n) {
          const int e = (m * NS + n) * kReactorBlock;
          const double jmn = rhoY[n] > 1e-30 ? a[e] / rhoY[n] * 1e+3 / tb->imw[m] : 0.0;
          a[e] = (m == n ? 1.0 : 0.0) - hs * kRkGamma * jmn;
        }
      lu_factor<NS, kReactorBlock>(a, ipiv);
#pragma unroll
      for (int l = 0; l < NS; ++l) k1[l] = omega[l];
      lu_apply<NS, kReactorBlock>(a, ipiv, k1);
      double imw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        ys[l] = rhoY[l] + hs * kRkA21 * k1[l];
        yk[l] = ys[l] / rho;
        imw += yk[l] * tb->imw[l];
      }
      double ts = reactor_temperature<M>(tb, t, imw, yk, e);
      reactor_source<M>(tb, rho, ts, yk, omega, nullptr);
#pragma unroll
      for (int l = 0; l < NS; ++l) k2[l] = omega[l] + kRkG21 * k1[l];
      lu_apply<NS, kReactorBlock>(a, ipiv, k2);
#pragma unroll
      for (int l = 0; l < NS; ++l) k2[l] -= kRkG21 * k1[l];
      imw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        ys[l] = rhoY[l] + hs * (kRkA31 * k1[l] + kRkA32 * k2[l]);
        yk[l] = ys[l] / rho;
        imw += yk[l] * tb->imw[l];
      }
      ts = reactor_temperature<M>(tb, ts, imw, yk, e);
      reactor_source<M>(tb, rho, ts, yk, omega, nullptr);
#pragma unroll
      for (int l = 0; l < NS; ++l) k3[l] = omega[l] + kRkG31 * k1[l] + kRkG32 * k2[l];
      lu_apply<NS, kReactorBlock>(a, ipiv, k3);
#pragma unroll
      for (int l = 0; l < NS; ++l) k3[l] -= kRkG31 * k1[l] + kRkG32 * k2[l];
#pragma unroll
      for (int l = 0; l < NS; ++l) k4[l] = omega[l] + kRkG41 * k1[l] + kRkG42 * k2[l] + kRkG43 * k3[l];
      lu_apply<NS, kReactorBlock>(a, ipiv, k4);
#pragma unroll
      for (int l = 0; l < NS; ++l) k4[l] -= kRkG41 * k1[l] + kRkG42 * k2[l] + kRkG43 * k3[l];
      if (adaptive) {
        double rn[NS], err = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          rn[l] = rhoY[l] + hs * (kRkB1 * k1[l] + kRkB2 * k2[l] + kRkB4 * k4[l]);
          double dy = hs * (kRkE1 * k1[l] + kRkE2 * k2[l] + kRkE3 * k3[l] + kRkE4 * k4[l]);
          dy /= rtol * rn[l] + atol;
          err += dy * dy;
        }
        err = sqrt(err / NS);
        double hstar = hs;
        if (steps > 0)
          hstar *= fmin(5.0, fmax(0.2, 0.8 * pow(err_last, 0.1) * pow(err + 1e-20, -0.175)));
        else if (err > 1)
          hstar *= 0.2;
        if (err <= 1) {
#pragma unroll
          for (int l = 0; l < NS; ++l) {
            rhoY[l] = rn[l];
            yk[l] = rn[l] / rho;
          }
          time += hs;
          err_last = err;
        } else {
          accepted = false;
          ++rejected;
#pragma unroll
          for (int l = 0; l < NS; ++l) yk[l] = rhoY[l] / rho;
        }
        h = hstar;
      } else {
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          rhoY[l] += hs * (kRkB1 * k1[l] + kRkB2 * k2[l] + kRkB4 * k4[l]);
          yk[l] = rhoY[l] / rho;
        }
        time += hs;
      }
    } else if (scheme == 3) {
      // SSP-RK3 (src/ConstVolumeReactor.cpp:708-745)
      double rn[NS], ys[NS], imw = 0;
      reactor_source<M>(tb, rho, t, yk, omega, nullptr);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rn[l] = rhoY[l] + omega[l] * h;
        ys[l] = rn[l] / rho;
        imw += ys[l] * tb->imw[l];
      }
      double ts = reactor_temperature<M>(tb, t, imw, ys, e);
      reactor_source<M>(tb, rho, ts, ys, omega, nullptr);
      imw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rn[l] = 0.75 * rhoY[l] + 0.25 * (rn[l] + omega[l] * h);
        ys[l] = rn[l] / rho;
        imw += ys[l] * tb->imw[l];
      }
      ts = reactor_temperature<M>(tb, ts, imw, ys, e);
      reactor_source<M>(tb, rho, ts, ys, omega, nullptr);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rhoY[l] = (rhoY[l] + 2.0 * (rn[l] + omega[l] * h)) / 3.0;
        if (rhoY[l] < 0) rhoY[l] = 0;
        yk[l] = rhoY[l] / rho;
      }
      time += h;
    } else {
      // point-implicit with the diagonal source Jacobian (src/ConstVolumeReactor.cpp:405-440, 1760-1799)
      double *zz = lu_lds + threadIdx.x;  // the source Jacobian, element stride kReactorBlock
      reactor_source<M, kReactorBlock>(tb, rho, t, yk, omega, zz);
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const double jmm = rhoY[m] > 1e-30 ? zz[(m * NS + m) * kReactorBlock] / rhoY[m] * 1e+3 / tb->imw[m] : 0.0;
        omega[m] /= 1 - h * jmm;
      }
      double sum = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rhoY[l] += omega[l] * h;
        if (rhoY[l] < 0) rhoY[l] = 0;
        sum += rhoY[l];
      }
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        if (conserve) {
          yk[l] = rhoY[l] / sum;
          rhoY[l] = yk[l] * rho;
        } else {
          yk[l] = rhoY[l] / rho;
        }
      }
      time += h;
    }
    if (!accepted) continue;
    double imw = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) imw += yk[l] * tb->imw[l];
    t = reactor_temperature<M>(tb, t, imw, yk, e);
    ++steps;
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) P.rhoY[l * cap + s] = rhoY[l];
  P.t[s] = t;
  P.time[s] = time;
  P.h[s] = h;
  P.err_last[s] = err_last;
  P.steps[s] = steps;
  P.rejected[s] = rejected;
  if (!fail) fail = reactor_fault<NS>(t, time, h, t_end, rhoY, steps + rejected, step_cap);
  if (fail)
    P.status[s] = kSlotFailed + fail;
  else if (time >= t_end)
    P.status[s] = 2;
}

// ---------------------------------------------------------------------------
// The same reactors with one wave per slot (latency): the state and the step
// control are held by every lane alike, and the costly parts of a step are
// spread over the lanes: G/RT, h and cp per species, the rates of progress per
// reaction, the creation / destruction sums per species and the source
// Jacobian per matrix element, the LU factorisation per row.  Every sum runs
// in the order of the one-lane kernel above, so a slot's result is the same
// bit for bit under either kernel.
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

// h and cp of species i at t (the arithmetic of thermo_h_cp)
template <int M>
DEV void thermo_h_cp1(const Tab *__restrict__ tb, int i, double t, double &h, double &cp) {
  if constexpr (kNasa9(M)) {
    nasa9_h_cp<true, true>(tb, i, t, h, cp);
    h *= tb->gas_const[i];
    cp *= tb->gas_const[i];
  } else {
    const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
    const double tl = tb->t_low[i], th = tb->t_high[i];
    double vh, vc;
    if (t < tl || t > th) {
      const double tt = t < tl ? tl : th;
      const double *c = t < tl ? tb->low_coeff[i] : tb->high_coeff[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      vh = nasa_h(c, tt, tt2, tt3, tt4, tt5);
      vc = nasa_cp(c, tt, tt2, tt3, tt4);
      vh += vc * (t - tt);
    } else {
      const bool lo = t < tb->t_mid[i];
      double c[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) c[k] = lo ? tb->low_coeff[i][k] : tb->high_coeff[i][k];
      vh = nasa_h(c, t, t2, t3, t4, t5);
      vc = nasa_cp(c, t, t2, t3, t4);
    }
    h = vh * tb->gas_const[i];
    cp = vc * tb->gas_const[i];
  }
}

template <int M>
struct ReactorWaveLds {
  static constexpr int NS = kNs(M);
  double c[NS], g[NS], h[NS], cp[NS], cr[NS], od[NS];
  double q1[COREFL_MAX_REAC], q2[COREFL_MAX_REAC];
  double a[NS * NS];  // source Jacobian, then the ROK4E iteration matrix and its LU factors
  // forward / backward stoichiometric coefficients (r, l) of the mechanism, staged once per launch
  double sf[COREFL_MAX_REAC * NS], sb[COREFL_MAX_REAC * NS];
};

// Stages the stoichiometry of the mechanism into the wave's LDS (the reaction
// lists' coefficients; a species takes part in reaction r when either is not 0).
// Read from LDS the species sums below take less time than walking the
// reaction lists in global memory.
template <int M>
DEV void wave_stage_stoichiometry(const Tab *__restrict__ tb, ReactorWaveLds<M> &L) {
  constexpr int NS = kNs(M);
  const int n = tb->n_reac * NS;
  for (int e = threadIdx.x; e < n; e += kWave) {
    const int r = e / NS, l = e % NS;
    L.sf[e] = static_cast<double>(tb->stoi_f[r][l]);
    L.sb[e] = static_cast<double>(tb->stoi_b[r][l]);
  }
  __syncthreads();
}

// v[i] of a register array for a lane-dependent i
template <int N>
DEV double pick(const double *v, int i) {
  double x = 0;
#pragma unroll
  for (int l = 0; l < N; ++l) x = l == i ? v[l] : x;
  return x;
}

// reactor_source over the wave: omega (every lane) and, with JAC, the source
// Jacobian sums zz in L.a
template <int M, bool JAC>
DEV void wave_source(const Tab *__restrict__ tb, ReactorWaveLds<M> &L, double rho, double t, const double *yk,
                     double *omega) {
  constexpr int NS = kNs(M);
  const int lane = threadIdx.x, nr = tb->n_reac;
  if (lane < NS) {
    L.c[lane] = rho * pick<NS>(yk, lane) * tb->imw[lane] * 1e-3;
    L.g[lane] = gibbs1<M>(tb, lane, t);
  }
  __syncthreads();
  if (lane < nr) {
    // concentrations and G/RT read from LDS at the reaction's own species indices
    const double ln_t = log(t), inv_t = 1.0 / t;
    double q1, q2;
    reaction_progress<M>(tb, lane, t, t, false, ln_t, inv_t, ln_std_conc(ln_t), L.c, L.g, q1, q2);
    L.q1[lane] = q1;
    L.q2[lane] = q2;
  }
  __syncthreads();
  if (lane < NS) {
    double cr = 0, od = 0;
#pragma unroll 4
    for (int r = 0; r < nr; ++r) {
      const double sf = L.sf[r * NS + lane], sb = L.sb[r * NS + lane];
      if (sf != 0.0 || sb != 0.0) {
        const double q1 = L.q1[r], q2 = L.q2[r];
        cr += q2 * sf + q1 * sb;
        od += q1 * sf + q2 * sb;
      }
    }
    L.cr[lane] = cr;
    L.od[lane] = od;
  }
  if constexpr (JAC) {
    for (int e = lane; e < NS * NS; e += kWave) {
      const int l = e / NS, l2 = e % NS;
      double zz = 0;
#pragma unroll 4
      for (int r = 0; r < nr; ++r) {
        const double sf = L.sf[r * NS + l], sb = L.sb[r * NS + l];
        const double sf2 = L.sf[r * NS + l2], sb2 = L.sb[r * NS + l2];
        if ((sf != 0.0 || sb != 0.0) && (sf2 != 0.0 || sb2 != 0.0))
          zz += (sb - sf) * (sf2 * L.q1[r] - sb2 * L.q2[r]);
      }
      L.a[e] = zz;
    }
  }
  __syncthreads();
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double conv = 1e+3 / tb->imw[l];
    omega[l] = L.cr[l] * conv - L.od[l] * conv;
  }
}

// reactor_temperature over the wave (h and cp per species on its lane)
template <int M>
DEV double wave_temperature(const Tab *__restrict__ tb, ReactorWaveLds<M> &L, double t0, double imw, const double *yk,
                            double e) {
  constexpr int NS = kNs(M);
  const int lane = threadIdx.x;
  double t = t0, err = 1e+6;
  const double r = kRu * imw;
  int iter = 0;
  while (err > 1e-8 && iter++ < 100) {
    if (lane < NS) {
      double h, cp;
      thermo_h_cp1<M>(tb, lane, t, h, cp);
      L.h[lane] = h;
      L.cp[lane] = cp;
    }
    __syncthreads();
    double cp_tot = 0, hm = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp_tot += L.cp[l] * yk[l];
      hm += L.h[l] * yk[l];
    }
    __syncthreads();  // read before the next iteration writes
    const double t1 = t - (hm - r * t - e) / (cp_tot - r);
    err = fabs(1 - t1 / t);
    t = t1;
  }
  return t;
}

// lu_factor over the wave: each step's column, row and scaling spread over the rows / columns
template <int NS>
DEV void wave_lu_factor(double *a, int *ipiv) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    if (lane >= n && lane < NS) {
      double v = a[lane * NS + n];
      for (int t = 0; t < n; ++t) v -= a[lane * NS + t] * a[t * NS + n];
      a[lane * NS + n] = v;
    }
    __syncthreads();
    int ik = n;
    double best = fabs(a[n * NS + n]);
#pragma unroll
    for (int m = n + 1; m < NS; ++m) {
      const double v = fabs(a[m * NS + n]);
      if (v > best) {
        best = v;
        ik = m;
      }
    }
    ipiv[n] = ik;
    __syncthreads();  // every lane has read column n
    if (ik != n && lane < NS) {
      const double x = a[n * NS + lane], w = a[ik * NS + lane];
      a[n * NS + lane] = w;
      a[ik * NS + lane] = x;
    }
    __syncthreads();
    if (lane > n && lane < NS) {
      double v = a[n * NS + lane];
      for (int t = 0; t < n; ++t) v -= a[n * NS + t] * a[t * NS + lane];
      a[n * NS + lane] = v;
      a[lane * NS + n] /= a[n * NS + n];
    }
    __syncthreads();
  }
}

template <int M>
__global__ __launch_bounds__(kWave) void k_reactors_wave(ReactorPool P, const Tab *__restrict__ tb, int max_steps,
                                                         int n_slots) {
  constexpr int NS = kNs(M);
  __shared__ ReactorWaveLds<M> L;
  const int s = blockIdx.x;
  if (s >= n_slots || P.status[s] != 1) return;
  const int lane = threadIdx.x;
  wave_stage_stoichiometry<M>(tb, L);
  const long cap = P.cap;
  double rhoY[NS], yk[NS], omega[NS];
  const double rho = P.rho[s], e = P.e[s], t_end = P.t_end[s], rtol = P.rtol[s], atol = P.atol[s];
  const int scheme = P.scheme[s], adaptive = P.adaptive[s], conserve = P.conserve[s];
  double t = P.t[s], time = P.time[s], h = P.h[s], err_last = P.err_last[s];
  long steps = P.steps[s], rejected = P.rejected[s];
  const long step_cap = P.max_steps[s];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    rhoY[l] = P.rhoY[l * cap + s];
    yk[l] = rhoY[l] / rho;
  }
  int fail = 0;
  for (int it = 0; it < max_steps; ++it) {
    fail = reactor_fault<NS>(t, time, h, t_end, rhoY, steps + rejected, step_cap);
    if (fail || time >= t_end) break;
    bool accepted = true;
    if (scheme == 1) {
      double k1[NS], k2[NS], k3[NS], k4[NS], ys[NS];
      int ipiv[NS];
      wave_source<M, true>(tb, L, rho, t, yk, omega);
      const double hs = fmin(h, t_end - time);
      for (int el = lane; el < NS * NS; el += kWave) {
        const int m = el / NS, n = el % NS;
        const double rn = pick<NS>(rhoY, n);
        const double jmn = rn > 1e-30 ? L.a[el] / rn * 1e+3 / tb->imw[m] : 0.0;
        L.a[el] = (m == n ? 1.0 : 0.0) - hs * kRkGamma * jmn;
      }
      __syncthreads();
      wave_lu_factor<NS>(L.a, ipiv);
#pragma unroll
      for (int l = 0; l < NS; ++l) k1[l] = omega[l];
      lu_apply<NS>(L.a, ipiv, k1);
      double imw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        ys[l] = rhoY[l] + hs * kRkA21 * k1[l];
        yk[l] = ys[l] / rho;
        imw += yk[l] * tb->imw[l];
      }
      double ts = wave_temperature<M>(tb, L, t, imw, yk, e);
      wave_source<M, false>(tb, L, rho, ts, yk, omega);
#pragma unroll
      for (int l = 0; l < NS; ++l) k2[l] = omega[l] + kRkG21 * k1[l];
      lu_apply<NS>(L.a, ipiv, k2);
#pragma unroll
      for (int l = 0; l < NS; ++l) k2[l] -= kRkG21 * k1[l];
      imw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        ys[l] = rhoY[l] + hs * (kRkA31 * k1[l] + kRkA32 * k2[l]);
        yk[l] = ys[l] / rho;
        imw += yk[l] * tb->imw[l];
      }
      ts = wave_temperature<M>(tb, L, ts, imw, yk, e);
      wave_source<M, false>(tb, L, rho, ts, yk, omega);
#pragma unroll
      for (int l = 0; l < NS; ++l) k3[l] = omega[l] + kRkG31 * k1[l] + kRkG32 * k2[l];
      lu_apply<NS>(L.a, ipiv, k3);
#pragma unroll
      for (int l = 0; l < NS; ++l) k3[l] -= kRkG31 * k1[l] + kRkG32 * k2[l];
#pragma unroll
      for (int l = 0; l < NS; ++l) k4[l] = omega[l] + kRkG41 * k1[l] + kRkG42 * k2[l] + kRkG43 * k3[l];
      lu_apply<NS>(L.a, ipiv, k4);
#pragma unroll
      for (int l = 0; l < NS; ++l) k4[l] -= kRkG41 * k1[l] + kRkG42 * k2[l] + kRkG43 * k3[l];
      __syncthreads();  // the factors are read before the next step's Jacobian overwrites them
      if (adaptive) {
        double rn[NS], err = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          rn[l] = rhoY[l] + hs * (kRkB1 * k1[l] + kRkB2 * k2[l] + kRkB4 * k4[l]);
          double dy = hs * (kRkE1 * k1[l] + kRkE2 * k2[l] + kRkE3 * k3[l] + kRkE4 * k4[l]);
          dy /= rtol * rn[l] + atol;
          err += dy * dy;
        }
        err = sqrt(err / NS);
        double hstar = hs;
        if (steps > 0)
          hstar *= fmin(5.0, fmax(0.2, 0.8 * pow(err_last, 0.1) * pow(err + 1e-20, -0.175)));
        else if (err > 1)
          hstar *= 0.2;
        if (err <= 1) {
#pragma unroll
          for (int l = 0; l < NS; ++l) {
            rhoY[l] = rn[l];
            yk[l] = rn[l] / rho;
          }
          time += hs;
          err_last = err;
        } else {
          accepted = false;
          ++rejected;
#pragma unroll
          for (int l = 0; l < NS; ++l) yk[l] = rhoY[l] / rho;
        }
        h = hstar;
      } else {
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          rhoY[l] += hs * (kRkB1 * k1[l] + kRkB2 * k2[l] + kRkB4 * k4[l]);
          yk[l] = rhoY[l] / rho;
        }
        time += hs;
      }
    } else if (scheme == 3) {
      double rn[NS], ys[NS], imw = 0;
      wave_source<M, false>(tb, L, rho, t, yk, omega);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rn[l] = rhoY[l] + omega[l] * h;
        ys[l] = rn[l] / rho;
        imw += ys[l] * tb->imw[l];
      }
      double ts = wave_temperature<M>(tb, L, t, imw, ys, e);
      wave_source<M, false>(tb, L, rho, ts, ys, omega);
      imw = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rn[l] = 0.75 * rhoY[l] + 0.25 * (rn[l] + omega[l] * h);
        ys[l] = rn[l] / rho;
        imw += ys[l] * tb->imw[l];
      }
      ts = wave_temperature<M>(tb, L, ts, imw, ys, e);
      wave_source<M, false>(tb, L, rho, ts, ys, omega);
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rhoY[l] = (rhoY[l] + 2.0 * (rn[l] + omega[l] * h)) / 3.0;
        if (rhoY[l] < 0) rhoY[l] = 0;
        yk[l] = rhoY[l] / rho;
      }
      time += h;
    } else {
      wave_source<M, true>(tb, L, rho, t, yk, omega);
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const double jmm = rhoY[m] > 1e-30 ? L.a[m * NS + m] / rhoY[m] * 1e+3 / tb->imw[m] : 0.0;
        omega[m] /= 1 - h * jmm;
      }
      __syncthreads();  // the diagonal is read before the next step's Jacobian overwrites it
      double sum = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rhoY[l] += omega[l] * h;
        if (rhoY[l] < 0) rhoY[l] = 0;
        sum += rhoY[l];
      }
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        if (conserve) {
          yk[l] = rhoY[l] / sum;
          rhoY[l] = yk[l] * rho;
        } else {
          yk[l] = rhoY[l] / rho;
        }
      }
      time += h;
    }
    if (!accepted) continue;
    double imw = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) imw += yk[l] * tb->imw[l];
    t = wave_temperature<M>(tb, L, t, imw, yk, e);
    ++steps;
  }
  if (lane != 0) return;
#pragma unroll
  for (int l = 0; l < NS; ++l) P.rhoY[l * cap + s] = rhoY[l];
  P.t[s] = t;
  P.time[s] = time;
  P.h[s] = h;
  P.err_last[s] = err_last;
  P.steps[s] = steps;
  P.rejected[s] = rejected;
  if (!fail) fail = reactor_fault<NS>(t, time, h, t_end, rhoY, steps + rejected, step_cap);
  if (fail)
    P.status[s] = kSlotFailed + fail;
  else if (time >= t_end)
    P.status[s] = 2;
}

// Admission: the request's state enters its slot; rho from p = rho R T and
// E = sum Y h(T) - R T (src/ConstVolumeReactor.cpp:300-334).
template <int M>
__global__ void k_reactor_admit(ReactorPool P, const Tab *__restrict__ tb, const double *req, int n_req) {
  constexpr int NS = kNs(M);
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_req) return;
  const double *q = req + static_cast<long>(r) * kReactorRecord;
  const int s = static_cast<int>(q[0]);
  if (s < 0 || s >= P.cap) return;
  const long cap = P.cap;
  const double t0 = q[1], p0 = q[2];
  double y[NS], h[NS], imw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    y[l] = q[11 + l];
    imw += y[l] * tb->imw[l];
  }
  thermo_h<M>(tb, t0, h);
  const double rho = p0 / (kRu * imw * t0);
  double e = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    e += y[l] * h[l];
    P.rhoY[l * cap + s] = rho * y[l];
  }
  e -= kRu * imw * t0;
  P.rho[s] = rho;
  P.e[s] = e;
  P.t[s] = t0;
  P.time[s] = 0.0;
  P.t_end[s] = q[3];
  P.h[s] = q[4];
  P.rtol[s] = q[5];
  P.atol[s] = q[6];
  P.scheme[s] = static_cast<int>(q[7]);
  P.adaptive[s] = static_cast<int>(q[8]);
  P.conserve[s] = static_cast<int>(q[9]);
  P.err_last[s] = 0.0;
  P.steps[s] = 0;
  P.rejected[s] = 0;
  P.max_steps[s] = static_cast<long>(q[10]);
  P.status[s] = 1;
}

// ---------------------------------------------------------------------------
// WENO5-JS / WENO7-Z (shock branch), src/WENO.cu:3033-3196
// ---------------------------------------------------------------------------
// The nonlinear weights d_k / (eps + beta_k)^2 are normalised by their sum,
// so they are scaled by prod_{m} (eps + beta_m)^2 first: one division per
// side instead of four.  (eps + beta) lies between eps >= 1e-40 (rho v)^2 and
// the square of the largest characteristic flux, so the products of two
// squared terms stay inside the double range.
DEV double weno5_side(double q0, double q1, double q2, double b0, double b1, double b2, double d0, double d1,
                      double d2, double eps) {
  const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  const double e12 = e1 * e2, e02 = e0 * e2, e01 = e0 * e1;
  const double a0 = d0 * (e12 * e12), a1 = d1 * (e02 * e02), a2 = d2 * (e01 * e01);
  return (a0 * q0 + a1 * q1 + a2 * q2) * fast_rcp(a0 + a1 + a2);
}

// The left-biased half of WENO5-JS at the face right of vp[2] (vp = the
// values at i-2 .. i+2) and the right-biased half at the face left of vm[2]
// (vm = the values at i-1 .. i+3), as the two halves of weno5 below.
DEV double weno5_left(const double *vp, double eps) {
  constexpr double one6th = 1.0 / 6, thirteen12th = 13.0 / 12;
  const double v0 = one6th * (2 * vp[2] + 5 * vp[3] - vp[4]);
  const double v1 = one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3]);
  const double v2 = one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2]);
  const double beta0 = thirteen12th * (vp[2] + vp[4] - 2 * vp[3]) * (vp[2] + vp[4] - 2 * vp[3]) +
                       0.25 * (3 * vp[2] - 4 * vp[3] + vp[4]) * (3 * vp[2] - 4 * vp[3] + vp[4]);
  const double beta1 = thirteen12th * (vp[1] + vp[3] - 2 * vp[2]) * (vp[1] + vp[3] - 2 * vp[2]) +
                       0.25 * (vp[1] - vp[3]) * (vp[1] - vp[3]);
  const double beta2 = thirteen12th * (vp[0] + vp[2] - 2 * vp[1]) * (vp[0] + vp[2] - 2 * vp[1]) +
                       0.25 * (vp[0] - 4 * vp[1] + 3 * vp[2]) * (vp[0] - 4 * vp[1] + 3 * vp[2]);
  return weno5_side(v0, v1, v2, beta0, beta1, beta2, 0.3, 0.6, 0.1, eps);
}

DEV double weno5_right(const double *vm, double eps) {
  constexpr double one6th = 1.0 / 6, thirteen12th = 13.0 / 12;
  const double v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
  const double v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
  const double v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
  const double beta0 = thirteen12th * (vm[2] + vm[4] - 2 * vm[3]) * (vm[2] + vm[4] - 2 * vm[3]) +
                       0.25 * (3 * vm[2] - 4 * vm[3] + vm[4]) * (3 * vm[2] - 4 * vm[3] + vm[4]);
  const double beta1 = thirteen12th * (vm[1] + vm[3] - 2 * vm[2]) * (vm[1] + vm[3] - 2 * vm[2]) +
                       0.25 * (vm[1] - vm[3]) * (vm[1] - vm[3]);
  const double beta2 = thirteen12th * (vm[0] + vm[2] - 2 * vm[1]) * (vm[0] + vm[2] - 2 * vm[1]) +
                       0.25 * (vm[0] - 4 * vm[1] + 3 * vm[2]) * (vm[0] - 4 * vm[1] + 3 * vm[2]);
  return weno5_side(v0, v1, v2, beta0, beta1, beta2, 0.1, 0.6, 0.3, eps);
}

DEV double weno5(const double *vp, const double *vm, double eps) {
  constexpr double one6th = 1.0 / 6, thirteen12th = 13.0 / 12;
  double v0 = one6th * (2 * vp[2] + 5 * vp[3] - vp[4]);
  double v1 = one6th * (-vp[1] + 5 * vp[2] + 2 * vp[3]);
  double v2 = one6th * (2 * vp[0] - 7 * vp[1] + 11 * vp[2]);
  double beta0 = thirteen12th * (vp[2] + vp[4] - 2 * vp[3]) * (vp[2] + vp[4] - 2 * vp[3]) +
                 0.25 * (3 * vp[2] - 4 * vp[3] + vp[4]) * (3 * vp[2] - 4 * vp[3] + vp[4]);
  double beta1 = thirteen12th * (vp[1] + vp[3] - 2 * vp[2]) * (vp[1] + vp[3] - 2 * vp[2]) +
                 0.25 * (vp[1] - vp[3]) * (vp[1] - vp[3]);
  double beta2 = thirteen12th * (vp[0] + vp[2] - 2 * vp[1]) * (vp[0] + vp[2] - 2 * vp[1]) +
                 0.25 * (vp[0] - 4 * vp[1] + 3 * vp[2]) * (vp[0] - 4 * vp[1] + 3 * vp[2]);
  const double fPlus = weno5_side(v0, v1, v2, beta0, beta1, beta2, 0.3, 0.6, 0.1, eps);
  v0 = one6th * (11 * vm[2] - 7 * vm[3] + 2 * vm[4]);
  v1 = one6th * (2 * vm[1] + 5 * vm[2] - vm[3]);
  v2 = one6th * (-vm[0] + 5 * vm[1] + 2 * vm[2]);
  beta0 = thirteen12th * (vm[2] + vm[4] - 2 * vm[3]) * (vm[2] + vm[4] - 2 * vm[3]) +
          0.25 * (3 * vm[2] - 4 * vm[3] + vm[4]) * (3 * vm[2] - 4 * vm[3] + vm[4]);
  beta1 = thirteen12th * (vm[1] + vm[3] - 2 * vm[2]) * (vm[1] + vm[3] - 2 * vm[2]) +
          0.25 * (vm[1] - vm[3]) * (vm[1] - vm[3]);
  beta2 = thirteen12th * (vm[0] + vm[2] - 2 * vm[1]) * (vm[0] + vm[2] - 2 * vm[1]) +
          0.25 * (vm[0] - 4 * vm[1] + 3 * vm[2]) * (vm[0] - 4 * vm[1] + 3 * vm[2]);
  const double fMinus = weno5_side(v0, v1, v2, beta0, beta1, beta2, 0.1, 0.6, 0.3, eps);
  return fPlus + fMinus;
}

DEV double w7beta(double s1, double s2, double s3) {
  constexpr double d12 = 13.0 / 12.0, d13 = 1043.0 / 960, d14 = 1.0 / 12;
  // s1^2 + d12 s2^2 + d13 s3^2 + d14 s1 s3
  return fma(d13 * s3, s3, fma(d12 * s2, s2, s1 * fma(d14, s3, s1)));
}

// One half of WENO7-Z on the upwind-ordered values v[0..6] (v[3] is the
// point left of the face).  The three derivative estimates s1, s2, s3 of
// each candidate stencil (the reference's polynomials, src/WENO.cu:3094-3117)
// are formed from the shared differences d_k = v_{k+1} - v_k, their second
// differences D_k and third differences T_k: e.g. the left stencil's
// (-2 v0 + 9 v1 - 18 v2 + 11 v3) / 6 is d2 + D1 / 2 + T0 / 3.
DEV double weno7_half(double v0_, double v1_, double v2_, double v3_, double v4_, double v5_, double v6_,
                      double eps) {
  constexpr double one12th = 1.0 / 12, third = 1.0 / 3, sixth = 1.0 / 6;
  constexpr double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35;
  const double d2 = v3_ - v2_, d3 = v4_ - v3_;
  const double D0 = fma(-2.0, v1_, v0_ + v2_), D1 = fma(-2.0, v2_, v1_ + v3_), D2 = fma(-2.0, v3_, v2_ + v4_),
               D3 = fma(-2.0, v4_, v3_ + v5_), D4 = fma(-2.0, v5_, v4_ + v6_);
  const double T0 = D1 - D0, T1 = D2 - D1, T2 = D3 - D2, T3 = D4 - D3;
  const double beta0 = w7beta(fma(third, T0, fma(0.5, D1, d2)), D1 + T0, T0);
  const double beta1 = w7beta(fma(-sixth, T1, fma(0.5, D2, d2)), D2, T1);
  const double beta2 = w7beta(fma(-sixth, T2, fma(-0.5, D2, d3)), D2, T2);
  const double beta3 = w7beta(fma(third, T3, fma(-0.5, D3, d3)), D3 - T3, T3);
  double tau = beta0 + 3 * beta1 - 3 * beta2 - beta3;
  tau *= tau;
  // 1/(eps + beta_k) from one division of the product of the four terms
  // (each lies between eps >= 1e-40 (rho v)^2 and the square of the largest
  // characteristic flux, so the product stays inside the double range).
  const double e0 = eps + beta0, e1 = eps + beta1, e2 = eps + beta2, e3 = eps + beta3;
  const double e01 = e0 * e1, e23 = e2 * e3;
  const double inv = fast_rcp(e01 * e23);
  const double r0 = inv * (e1 * e23), r1 = inv * (e0 * e23), r2 = inv * (e01 * e3), r3 = inv * (e01 * e2);
  const double a0 = c0 + c0 * tau * (r0 * r0);
  const double a1 = c1 + c1 * tau * (r1 * r1);
  const double a2 = c2 + c2 * tau * (r2 * r2);
  const double a3 = c3 + c3 * tau * (r3 * r3);
  const double q0 = -3 * v0_ + 13 * v1_ - 23 * v2_ + 25 * v3_;
  const double q1 = v1_ - 5 * v2_ + 13 * v3_ + 3 * v4_;
  const double q2 = -v2_ + 7 * v3_ + 7 * v4_ - v5_;
  const double q3 = 3 * v3_ + 13 * v4_ - 5 * v5_ + v6_;
  return one12th * (a0 * q0 + a1 * q1 + a2 * q2 + a3 * q3) * fast_rcp(a0 + a1 + a2 + a3);
}

DEV double weno7(const double *vp, const double *vm, double eps) {
  const double fp = weno7_half(vp[0], vp[1], vp[2], vp[3], vp[4], vp[5], vp[6], eps);
  // The minus side starts after the plus side: evaluating both at once
  // doubles the live registers and halves the waves per SIMD.
  double m0 = vm[0], m1 = vm[1], m2 = vm[2], m3 = vm[3], m4 = vm[4], m5 = vm[5], m6 = vm[6];
  asm volatile("" : "+v"(m0), "+v"(m1), "+v"(m2), "+v"(m3), "+v"(m4), "+v"(m5), "+v"(m6) : "v"(fp));
  const double fm = weno7_half(m6, m5, m4, m3, m2, m1, m0, eps);
  return fp + fm;
}

// The linear (upwind UD5 / UD7) branches of WENO5 / WENO7 that the hybrid
// scheme takes away from shocks (if_shock false), src/WENO.cu:3075-3087, 3177-3195.
DEV double weno5_linear(const double *vp, const double *vm) {
  constexpr double one6th = 1.0 / 6;
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

DEV double weno7_linear(const double *vp, const double *vm) {
  constexpr double c0 = 1.0 / 35, c1 = 12.0 / 35, c2 = 18.0 / 35, c3 = 4.0 / 35, one12th = 1.0 / 12;
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

template <bool W7>
DEV double weno_flux(const double *vp, const double *vm, double eps, bool shock) {
  if constexpr (W7)
    return shock ? weno7(vp, vm, eps) : weno7_linear(vp, vm);
  else
    return shock ? weno5(vp, vm, eps) : weno5_linear(vp, vm);
}

// Options of the WENO kernels: the hybrid switch and the species positivity
// limiter (src/WENO.cu:839-849, 991-1085, 1459-1527).
struct WenoOpt {
  const double *sensor;  // shock sensor plane (padded) when the hybrid threshold is set, else nullptr
  double threshold;
  int pp;                // positivity limiter on
  double pp_dt;          // DParameter::dt (fixed step) or 0: then dt_local of the face's left node
  double alpha;          // 1/3 in 3-D, 1/2 in 2-D
  double rho_ref, v_ref;
};

// if_shock of the face whose left node is pf: a sensor value above the
// threshold anywhere in [left - ngg + 1, left + ngg]; sensor ghosts stay 0.
DEV bool face_shock(const Zone &z, const WenoOpt &o, long pf, long stride) {
  if (!o.sensor) return true;
  bool shock = false;
  for (int ii = -z.ngg + 1; ii <= z.ngg; ++ii)
    if (o.sensor[pf + ii * stride] > o.threshold) shock = true;
  return shock;
}

// dt_local read with the interior array's flat index; face -1 of the first
// row (a read before the array in the reference) is clamped to element 0.
DEV double pp_time_step(const Zone &z, const WenoOpt &o, int i, int j, int k) {
  if (o.pp_dt > 0) return o.pp_dt;
  const long q = z.dx(i, j, k);
  return z.dt_local[q < 0 ? 0 : q];
}

// One scalar flux through the positivity limiter, src/WENO.cu:1047-1083:
// blended with the first-order flux 0.5 f1_left + 0.5 f1_right where
// rhoY_left - dt F or rhoY_right + dt F would turn negative.
DEV double pp_limit(double fc, double alphaL, double alphaR, double rhoYL, double rhoYR, double dt, double f1_left,
                    double f1_right, double rho_ref) {
  double f1 = 0.0, theta_p = 1.0, theta_m = 1.0;
  bool f1_computed = false;
  const double up = alphaL * rhoYL - dt * fc;
  if (up < 0) {
    f1 = 0.5 * f1_left + 0.5 * f1_right;
    f1_computed = true;
    const double up_lf = alphaL * rhoYL - dt * f1;
    if (fabs(up - up_lf) > 1e-20 * rho_ref) theta_p = fmin(fmax((0 - up_lf) / (up - up_lf), 0.0), 1.0);
  }
  const double um = alphaR * rhoYR + dt * fc;
  if (um < 0) {
    if (!f1_computed) f1 = 0.5 * f1_left + 0.5 * f1_right;
    const double um_lf = alphaR * rhoYR + dt * f1;
    if (fabs(um - um_lf) > 1e-20 * rho_ref) theta_m = fmin(fmax((0 - um_lf) / (um - um_lf), 0.0), 1.0);
  }
  return fmin(theta_p, theta_m) * (fc - f1) + f1;
}

// ---------------------------------------------------------------------------
// Characteristic WENO flux + fused flux difference.
//
// One thread per face; a block covers TS consecutive faces along the sweep
// and TX lateral lines.  The stencil points of all faces are staged in LDS
// (Q = cv[nv], F0 = U_k, F1..F3 = rho u_i U_k + p k_i, F4 = p, F5 = c|grad k|,
// F6 = J, as in src/WENO.cu:1096-1138).  After the face fluxes are computed
// they are exchanged through LDS and each block updates the TS-1 points whose
// two faces it owns: dq -= F(i+1/2) - F(i-1/2) (src/WENO.cu:1532-1559).
// Consecutive blocks overlap by one face.
//
// Face metrics average (i,j,k) and (i+1,j,k) in every direction, as the
// reference's kernels do (src/WENO.cu:886-888, 1615-1617, 2280-2282).
// ---------------------------------------------------------------------------
// Workgroup w runs on XCD w % 8 (round-robin dispatch).  Renumbering so that
// each XCD takes a contiguous range of logical workgroups keeps the rows a
// stencil shares (j +- 4 of the same plane) in one XCD's L2.
DEV long xcd_block(long w, long nblocks) {
  const long q = nblocks / 8, r = nblocks % 8, x = w % 8;
  return x * q + (x < r ? x : r) + w / 8;
}

// The tile of a 3-D sweep grid this workgroup takes: the linear workgroup
// index renumbered by xcd_block and decoded back, so that tiles adjacent in i
// (which share 128-byte lines when a tile is 8 points wide) and in the sweep
// direction (which share halo rows) run in one XCD's L2.  The sweep direction
// is decoded right after i: for a z sweep the tiles along k come before those
// along j (with j first, the k neighbours lie a whole plane of tiles apart and
// their halo planes were fetched again, 1.3x the x sweep's bytes).
template <int DIR>
DEV void sweep_tile(int &bx, int &by, int &bz) {
  const long gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const long w = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * static_cast<long>(blockIdx.z)), gx * gy * gz);
  bx = static_cast<int>(w % gx);
  const long r = w / gx;
  if constexpr (DIR == 2) {
    bz = static_cast<int>(r % gz);
    by = static_cast<int>(r / gz);
  } else {
    by = static_cast<int>(r % gy);
    bz = static_cast<int>(r / gy);
  }
}

template <int M, bool W7, int DIR, int TX, int TS>
__global__ __launch_bounds__(TX *TS) void k_weno(Zone z, const Tab *__restrict__ tb, double eps_ref, WenoOpt o) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  constexpr int NSC = kNsc(M);  // transported scalars: species (and E_ve)
  constexpr int NV = kNv(M);
  constexpr int LO = W7 ? 3 : 2, HI = W7 ? 4 : 3;
  constexpr int NPT = LO + HI + 1;          // stencil points per face
  constexpr int NP = TS + LO + HI;          // staged positions along the sweep
  constexpr int NQ = NV + 9;                // staged values per point
  constexpr int NT = TX * TS;               // threads per block
  // Stencil staging.  The face fluxes for the difference reuse planes that
  // are no longer read: species fluxes go to the planes of rho u, rho v,
  // rho w, rho E and F1..F5 once the acoustic/shear rows are done, the other
  // five components to the planes of rho and the first four species once the
  // species rows are done.
  static_assert(NSC <= 9, "scalar flux slots cover at most nine scalars");
  constexpr int PL = NP * TX;  // one staged plane
  __shared__ double s[NQ * PL];
#define SQ(v, ps, tx) s[((v) * NP + (ps)) * TX + (tx)]
  auto flux_plane = [](int c) { return c < 5 ? (c == 0 ? 0 : 4 + c) : (c < 9 ? c - 4 : NV + c - 8); };

  const int tid = threadIdx.x;
  int tx, ts;
  // one line of x faces per block: faces along the lanes; otherwise the lateral line is fastest, so that
  // a wave's LDS reads are consecutive doubles
  if constexpr (DIR == 0 && TX == 1) {
    ts = tid;
    tx = 0;
  } else {
    tx = tid % TX;
    ts = tid / TX;
  }
  const int me = ts * TX + tx;  // slot of this thread in the per-thread LDS arrays
  const int n_sweep = DIR == 0 ? z.mx : (DIR == 1 ? z.my : z.mz);
  int f0, li, lj, lk;
  int bx, by, bz;
  sweep_tile<DIR>(bx, by, bz);
  if constexpr (DIR == 0) {
    f0 = bx * (TS - 1);
    lj = by * TX;
    lk = bz;
    li = 0;
  } else if constexpr (DIR == 1) {
    li = bx * TX;
    f0 = by * (TS - 1);
    lk = bz;
    lj = 0;
  } else {
    li = bx * TX;
    lj = by;
    f0 = bz * (TS - 1);
    lk = 0;
  }
  const long sz = z.sz;

  // ---- stage stencil points ----
  for (int e = tid; e < NP * TX; e += NT) {
    int ps, ex;
    if constexpr (DIR == 0) {
      ps = e % NP;
      ex = e / NP;
    } else {
      ex = e % TX;
      ps = e / TX;
    }
    const int pos = f0 - 1 - LO + ps;
    int i, j, k;
    bool lat_ok;
    if constexpr (DIR == 0) {
      i = pos;
      j = lj + ex;
      k = lk;
      lat_ok = j < z.my;
    } else if constexpr (DIR == 1) {
      i = li + ex;
      j = pos;
      k = lk;
      lat_ok = i < z.mx;
    } else {
      i = li + ex;
      j = lj;
      k = pos;
      lat_ok = i < z.mx;
    }
    if (!lat_ok || pos > n_sweep - 1 + HI) continue;
    const long p = z.ix(i, j, k);
    const double q0 = z.cv[p], q1 = z.cv[p + sz], q2 = z.cv[p + 2 * sz], q3 = z.cv[p + 3 * sz];
    SQ(0, ps, ex) = q0;
    SQ(1, ps, ex) = q1;
    SQ(2, ps, ex) = q2;
    SQ(3, ps, ex) = q3;
#pragma unroll
    for (int l = 4; l < NV; ++l) SQ(l, ps, ex) = z.cv[p + l * sz];
    const double m0 = z.metric[p + (3 * DIR) * sz];
    const double m1 = z.metric[p + (3 * DIR + 1) * sz];
    const double m2 = z.metric[p + (3 * DIR + 2) * sz];
    const double Uk = (q1 * m0 + q2 * m1 + q3 * m2) / q0;
    const double pk = z.bv[p + 4 * sz];
    SQ(NV + 7, ps, ex) = q0 * Uk;                  // rho U_k
    SQ(NV + 8, ps, ex) = (SQ(4, ps, ex) + pk) * Uk;  // (rho E + p) U_k
    SQ(NV + 0, ps, ex) = Uk;
    SQ(NV + 1, ps, ex) = fma(q1, Uk, pk * m0);
    SQ(NV + 2, ps, ex) = fma(q2, Uk, pk * m1);
    SQ(NV + 3, ps, ex) = fma(q3, Uk, pk * m2);
    SQ(NV + 4, ps, ex) = pk;
    double cgk = norm3(m0, m1, m2);
    if constexpr (NS > 0)
      cgk *= z.acoustic[p];
    else
      cgk *= sqrt(kGammaAir * kRair * z.bv[p + 5 * sz]);
    SQ(NV + 5, ps, ex) = cgk;
    SQ(NV + 6, ps, ex) = 0.5 * z.jac[p];  // half the Jacobian: every use takes J / 2
  }
  __syncthreads();

  // ---- face flux ----
  const int face = f0 - 1 + ts;
  int fi, fj, fk;
  bool lat_ok;
  if constexpr (DIR == 0) {
    fi = face;
    fj = lj + tx;
    fk = lk;
    lat_ok = fj < z.my;
  } else if constexpr (DIR == 1) {
    fi = li + tx;
    fj = face;
    fk = lk;
    lat_ok = fi < z.mx;
  } else {
    fi = li + tx;
    fj = lj;
    fk = face;
    lat_ok = fi < z.mx;
  }
  const bool face_ok = lat_ok && face <= n_sweep - 1;
  const int ic = ts + LO;
  const int base = ts;
  // carried from the acoustic/shear rows to the species rows
  double fc[5] = {0, 0, 0, 0, 0};
  // gamma R_l T - (gamma - 1) h_l of the Roe state per species: the scalar coefficients' numerators,
  // kept for the energy row's species terms
  [[maybe_unused]] double numr[NS > 0 ? NS : 1];
  double rlc = 0, rrc = 0, max_lambda = 0, t1b = 0, gm1 = 0, c2 = 1, inv_gm1 = 0, gamma = 0, t_eval = 1;
  bool shock = true;
  // c|grad k| of the face's two nodes for the positivity limiter (x only); its
  // staged plane is reused for fluxes before the species rows
  [[maybe_unused]] double cgk_l = 0, cgk_r = 0;
  if (face_ok) {
    const long pf = z.ix(fi, fj, fk), pm = z.ix(fi + 1, fj, fk);
    shock = face_shock(z, o, pf, DIR == 0 ? 1 : (DIR == 1 ? z.sx : z.sxy));
    if constexpr (DIR == 0 && NS > 0) {
      cgk_l = SQ(NV + 5, ic, tx);
      cgk_r = SQ(NV + 5, ic + 1, tx);
    }
    double kx = 0.5 * (z.metric[pf + (3 * DIR) * sz] + z.metric[pm + (3 * DIR) * sz]);
    double ky = 0.5 * (z.metric[pf + (3 * DIR + 1) * sz] + z.metric[pm + (3 * DIR + 1) * sz]);
    double kz = 0.5 * (z.metric[pf + (3 * DIR + 2) * sz] + z.metric[pm + (3 * DIR + 2) * sz]);

    const double rho_l = SQ(0, ic, tx), rho_r = SQ(0, ic + 1, tx);
    double temp1 = sqrt(rho_l * rho_r);
    rlc = 1 / (rho_l + temp1);
    rrc = 1 / (temp1 + rho_r);
    const double um = rlc * SQ(1, ic, tx) + rrc * SQ(1, ic + 1, tx);
    const double vm = rlc * SQ(2, ic, tx) + rrc * SQ(2, ic + 1, tx);
    const double wm = rlc * SQ(3, ic, tx) + rrc * SQ(3, ic + 1, tx);
    double cm;
    double sbq[NPT], sbf[NPT];
#pragma unroll
    for (int m = 0; m < NPT; ++m) sbq[m] = sbf[m] = 0;
    if constexpr (NS > 0) {
      // Roe-averaged mass fractions and the mixture thermodynamics at the Roe
      // temperature (compute_mixture_characteristic_thermo, src/Thermo.cuh:49-110)
      double svm[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) svm[l] = rlc * SQ(5 + l, ic, tx) + rrc * SQ(5 + l, ic + 1, tx);
      double r_mix_guess = 0.0;
#pragma unroll
      for (int l = 0; l < NS; ++l) r_mix_guess += svm[l] * tb->gas_const[l];
      const double troe = (rlc * SQ(NV + 4, ic, tx) + rrc * SQ(NV + 4, ic + 1, tx)) / fmax(r_mix_guess, 1e-12);
      t_eval = fmax(troe, 1.0);
      double h[NS], cp[NS];
      thermo_h_cp<M>(tb, t_eval, h, cp);
      if constexpr (kTT(M)) {
        // translational-rotational part: E_ve(T) and its c_v removed
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          double cv_ve;
          h[l] -= ve_energy1<true>(tb, l, t_eval, &cv_ve);
          cp[l] -= cv_ve;
        }
      }
      double r_local = 0, cp_local = 0, cv_local = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        r_local += svm[l] * tb->gas_const[l];
        cp_local += svm[l] * cp[l];
        cv_local += svm[l] * (cp[l] - tb->gas_const[l]);
      }
      gamma = cp_local / fmax(cv_local, 1e-8);
      cm = sqrt(fmax(gamma * r_local * t_eval, 1e-12));
      gm1 = gamma - 1.0;
      c2 = fmax(cm * cm, 1e-30);
      inv_gm1 = 1.0 / fmax(gm1, 1e-12);
      const double inv_c2 = 1.0 / c2;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const double num = gamma * tb->gas_const[l] * t_eval - gm1 * h[l];
        numr[l] = num;
        const double a = num * inv_c2;
#pragma unroll
        for (int m = 0; m < NPT; ++m) {
          const double qv = SQ(5 + l, base + m, tx);
          sbq[m] = fma(a, qv, sbq[m]);
          sbf[m] = fma(a, qv * SQ(NV + 0, base + m, tx), sbf[m]);
        }
      }
      if constexpr (kTT(M)) {
        // the E_ve row of the scalar coefficients: alpha = -(gamma-1)/c^2
        const double a = -gm1 * inv_c2;
#pragma unroll
        for (int m = 0; m < NPT; ++m) {
          const double qv = SQ(5 + NS, base + m, tx);
          sbq[m] = fma(a, qv, sbq[m]);
          sbf[m] = fma(a, qv * SQ(NV + 0, base + m, tx), sbf[m]);
        }
      }
    } else {
      const double troe = (rlc * SQ(NV + 4, ic, tx) + rrc * SQ(NV + 4, ic + 1, tx)) / kRair;
      cm = sqrt(kGammaAir * kRair * troe);
      gm1 = kGammaAir - 1;
    }
    temp1 = 1.0 / norm3(kx, ky, kz);
    kx *= temp1;
    ky *= temp1;
    kz *= temp1;
    const double Uk_bar = kx * um + ky * vm + kz * wm;
    double nx = 0, ny = 0, nz = 0;
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
    double qx = ky * nz - kz * ny, qy = kz * nx - kx * nz, qz = kx * ny - ky * nx;
    temp1 = 1.0 / norm3(qx, qy, qz);
    qx *= temp1;
    qy *= temp1;
    qz *= temp1;
    const double Un = nx * um + ny * vm + nz * wm;
    const double Uq = qx * um + qy * vm + qz * wm;
    const double temp2 = 1.0 / (cm * cm);
    const double temp3 = 0.5 * gm1 * (um * um + vm * vm + wm * wm);

#pragma unroll
    for (int m = 0; m < NPT; ++m) {
      const double lam = fabs(SQ(NV + 0, base + m, tx)) + SQ(NV + 5, base + m, tx);
      if (lam > max_lambda) max_lambda = lam;
    }

    // acoustic / entropy / shear rows of the left eigenvector matrix
    double fch0 = 0, fch1 = 0, fch2 = 0, fch3 = 0, fch4 = 0;
#pragma unroll 1
    for (int l = 0; l < 5; ++l) {
      double tf = 0.5;
      double L0, L1, L2, L3, L4;
      if (l == 0) {
        L0 = (temp3 + Uk_bar * cm) * temp2 * 0.5;
        L1 = -(gm1 * um + kx * cm) * temp2 * 0.5;
        L2 = -(gm1 * vm + ky * cm) * temp2 * 0.5;
        L3 = -(gm1 * wm + kz * cm) * temp2 * 0.5;
        L4 = gm1 * temp2 * 0.5;
      } else if (l == 1) {
        tf = 0;
        L0 = -Un;
        L1 = nx;
        L2 = ny;
        L3 = nz;
        L4 = 0;
      } else if (l == 2) {
        tf = 0;
        L0 = -Uq;
        L1 = qx;
        L2 = qy;
        L3 = qz;
        L4 = 0;
      } else if (l == 3) {
        tf = -1;
        L0 = 1 - temp3 * temp2;
        L1 = gm1 * um * temp2;
        L2 = gm1 * vm * temp2;
        L3 = gm1 * wm * temp2;
        L4 = -gm1 * temp2;
      } else {
        L0 = (temp3 - Uk_bar * cm) * temp2 * 0.5;
        L1 = -(gm1 * um - kx * cm) * temp2 * 0.5;
        L2 = -(gm1 * vm - ky * cm) * temp2 * 0.5;
        L3 = -(gm1 * wm - kz * cm) * temp2 * 0.5;
        L4 = gm1 * temp2 * 0.5;
      }
      double vP[NPT - 1], vM[NPT - 1];
      // Re-read the stencil from LDS for every row instead of keeping all of it
      // live in VGPRs across the row loop (the opaque index blocks hoisting).
      int row_base = base;
      asm volatile("" : "+v"(row_base));
#pragma unroll
      for (int m = 0; m < NPT; ++m) {
        const int ps = row_base + m;
        const double Q0 = SQ(0, ps, tx), Q1 = SQ(1, ps, tx), Q2 = SQ(2, ps, tx), Q3 = SQ(3, ps, tx),
                     Q4 = SQ(4, ps, tx);
        const double F1 = SQ(NV + 1, ps, tx), F2 = SQ(NV + 2, ps, tx), F3 = SQ(NV + 3, ps, tx),
                     hJ = SQ(NV + 6, ps, tx), FR = SQ(NV + 7, ps, tx), FE = SQ(NV + 8, ps, tx);
        double LFm = L0 * FR + L1 * F1 + L2 * F2 + L3 * F3 + L4 * FE;
        double LQm = L0 * Q0 + L1 * Q1 + L2 * Q2 + L3 * Q3 + L4 * Q4;
        if constexpr (NS > 0) {
          LFm += tf * sbf[m];
          LQm += tf * sbq[m];
        }
        if (m < NPT - 1) vP[m] = hJ * (LFm + max_lambda * LQm);
        if (m > 0) vM[m - 1] = hJ * (LFm - max_lambda * LQm);
      }
      const double fv = weno_flux<W7>(vP, vM, eps_ref, shock);
      if (l == 0)
        fch0 = fv;
      else if (l == 1)
        fch1 = fv;
      else if (l == 2)
        fch2 = fv;
      else if (l == 3)
        fch3 = fv;
      else
        fch4 = fv;
    }
    // back-projection, src/WENO.cu:1439-1457
    t1b = fch0 + fch3 + fch4;
    const double t3b = fch0 - fch4;
    fc[0] = t1b;
    fc[1] = um * t1b - kx * cm * t3b + nx * fch1 + qx * fch2;
    fc[2] = vm * t1b - ky * cm * t3b + ny * fch1 + qy * fch2;
    fc[3] = wm * t1b - kz * cm * t3b + nz * fch1 + qz * fch2;
    const double Hroe =
        rlc * (SQ(4, ic, tx) + SQ(NV + 4, ic, tx)) + rrc * (SQ(4, ic + 1, tx) + SQ(NV + 4, ic + 1, tx));
    fc[4] = Hroe * t1b - Uk_bar * cm * t3b + Un * fch1 + Uq * fch2 - cm * cm / gm1 * fch3;
  }
  if constexpr (NS > 0) {
    __syncthreads();  // rho u .. rho E and F1..F5 are no longer read
    if (face_ok) {
      // species rows: L.Q = rhoY_l - Y~_l rho, L.F = U_k L.Q
      double t2e = 0;
      // each species' numerator waits in its own face-flux slot until its row overwrites it
#pragma unroll
      for (int l = 0; l < NS; ++l) s[flux_plane(5 + l) * PL + me] = numr[l];
      // the split of a species row is L.Q (U_k +- lambda) J / 2 at each stencil point
      double ap[NPT], am[NPT];
#pragma unroll
      for (int m = 0; m < NPT; ++m) {
        const double hJ = SQ(NV + 6, base + m, tx), uk = SQ(NV + 0, base + m, tx);
        ap[m] = hJ * (uk + max_lambda);
        am[m] = hJ * (uk - max_lambda);
      }
#pragma unroll 1
      for (int l = 0; l < NSC; ++l) {
        const double num_l = l < NS ? s[flux_plane(5 + l) * PL + me] : 0.0;
        const double sv_l = rlc * SQ(5 + l, ic, tx) + rrc * SQ(5 + l, ic + 1, tx);
        double vP[NPT - 1], vM[NPT - 1];
#pragma unroll
        for (int m = 0; m < NPT; ++m) {
          const int ps = base + m;
          const double LQm = SQ(5 + l, ps, tx) - sv_l * SQ(0, ps, tx);
          if (m < NPT - 1) vP[m] = ap[m] * LQm;
          if (m > 0) vM[m - 1] = am[m] * LQm;
        }
        const double fv = weno_flux<W7>(vP, vM, eps_ref, shock);
        double fcl = sv_l * t1b + fv;
        if constexpr (DIR == 0) {
          // positivity limiter of the x kernel, src/WENO.cu:1459-1527 (the y and z
          // characteristic kernels have none)
          if (o.pp) {
            const double rhoYL = SQ(5 + l, ic, tx), rhoYR = SQ(5 + l, ic + 1, tx);
            const double UL = SQ(NV + 0, ic, tx), UR = SQ(NV + 0, ic + 1, tx);
            const double JL = 2.0 * SQ(NV + 6, ic, tx), JR = 2.0 * SQ(NV + 6, ic + 1, tx);
            const double fpl = 0.5 * JL * (UL + cgk_l) * rhoYL, fmr = 0.5 * JR * (UR - cgk_r) * rhoYR;
            fcl = pp_limit(fcl, 0.5 * o.alpha * JL, 0.5 * o.alpha * JR, rhoYL, rhoYR, pp_time_step(z, o, fi, fj, fk),
                           fpl + UL * JL * rhoYL, fmr - UR * JR * rhoYR, o.rho_ref);
          }
        }
        s[flux_plane(5 + l) * PL + me] = fcl;
        double ecoef;
        if (kTT(M) && l == NS) {
          ecoef = 1.0;  // E_ve row
        } else {
          // -alpha_l c^2 / (gamma - 1) with alpha_l = (gamma R_l T - (gamma - 1) h_l) / c^2
          ecoef = -num_l * inv_gm1;
        }
        t2e += ecoef * fv;
      }
      fc[4] += t2e;
    }
  }
  __syncthreads();  // every stencil read is done
#pragma unroll
  for (int c = 0; c < 5; ++c) s[flux_plane(c) * PL + me] = fc[c];
  __syncthreads();
  // ---- flux difference: dq -= F(i+1/2) - F(i-1/2), src/WENO.cu:1532-1559 ----
  if (face_ok && ts >= 1) {
    const long q = z.dx(fi, fj, fk);
    const long dsz = z.dsz;
    const int left = (ts - 1) * TX + tx;
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const double *fl = s + flux_plane(l) * PL;
      z.dq[q + l * dsz] -= fl[me] - fl[left];
    }
  }
#undef SQ
}

// ---------------------------------------------------------------------------
// Component-wise WENO (inviscid_scheme 51 / 71), src/WENO.cu:890-1086 (y
// 1619-1816, z 2284-2481): Lax-Friedrichs splitting of each conservative flux
// with the node's own |U_k| + c|grad k|, WENO5-JS / WENO7-Z (or their linear
// weights away from shocks) per component, the species positivity limiter,
// then dq -= F(i+1/2) - F(i-1/2).  Block decomposition as k_weno: TS faces
// along the sweep, TX lateral lines, one face of overlap; the split fluxes of
// the stencil nodes are staged in LDS and the face fluxes exchanged there.
// ---------------------------------------------------------------------------
template <int M, bool W7, int DIR, int TX, int TS>
__global__ __launch_bounds__(TX *TS) void k_weno_cp(Zone z, double eps_ref, WenoOpt o) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  constexpr int NV = kNv(M);
  constexpr int LO = W7 ? 3 : 2, HI = W7 ? 4 : 3;
  constexpr int NP = TS + LO + HI;
  constexpr int NT = TX * TS;
  constexpr int PL = NP * TX;
  static_assert(NV * NT <= 2 * NV * PL, "face fluxes fit in the staging buffer");
  __shared__ double s[2 * NV * PL];
#define FP(v, ps, tx) s[((v) * NP + (ps)) * TX + (tx)]
#define FM(v, ps, tx) s[((NV + (v)) * NP + (ps)) * TX + (tx)]
  const int tid = threadIdx.x;
  int tx, ts;
  if constexpr (DIR == 0) {
    ts = tid % TS;
    tx = tid / TS;
  } else {
    tx = tid % TX;
    ts = tid / TX;
  }
  const int me = ts * TX + tx;
  const int n_sweep = DIR == 0 ? z.mx : (DIR == 1 ? z.my : z.mz);
  int f0, li, lj, lk;
  int bx, by, bz;
  sweep_tile<DIR>(bx, by, bz);
  if constexpr (DIR == 0) {
    f0 = bx * (TS - 1);
    lj = by * TX;
    lk = bz;
    li = 0;
  } else if constexpr (DIR == 1) {
    li = bx * TX;
    f0 = by * (TS - 1);
    lk = bz;
    lj = 0;
  } else {
    li = bx * TX;
    lj = by;
    f0 = bz * (TS - 1);
    lk = 0;
  }
  const long sz = z.sz;
  const long stride = DIR == 0 ? 1 : (DIR == 1 ? z.sx : z.sxy);

  // ---- split fluxes of the stencil nodes ----
  for (int e = tid; e < NP * TX; e += NT) {
    int ps, ex;
    if constexpr (DIR == 0) {
      ps = e % NP;
      ex = e / NP;
    } else {
      ex = e % TX;
      ps = e / TX;
    }
    const int pos = f0 - 1 - LO + ps;
    int i, j, k;
    bool lat_ok;
    if constexpr (DIR == 0) {
      i = pos;
      j = lj + ex;
      k = lk;
      lat_ok = j < z.my;
    } else if constexpr (DIR == 1) {
      i = li + ex;
      j = pos;
      k = lk;
      lat_ok = i < z.mx;
    } else {
      i = li + ex;
      j = lj;
      k = pos;
      lat_ok = i < z.mx;
    }
    if (!lat_ok || pos > n_sweep - 1 + HI) continue;
    const long p = z.ix(i, j, k);
    const double q0 = z.cv[p], q1 = z.cv[p + sz], q2 = z.cv[p + 2 * sz], q3 = z.cv[p + 3 * sz],
                 q4 = z.cv[p + 4 * sz];
    const double m0 = z.metric[p + (3 * DIR) * sz], m1 = z.metric[p + (3 * DIR + 1) * sz],
                 m2 = z.metric[p + (3 * DIR + 2) * sz];
    const double Uk = (q1 * m0 + q2 * m1 + q3 * m2) / q0;
    double cgk = norm3(m0, m1, m2);
    if constexpr (NS > 0)
      cgk *= z.acoustic[p];
    else
      cgk *= sqrt(kGammaAir * kRair * z.bv[p + 5 * sz]);
    const double lambda0 = fabs(Uk) + cgk;
    const double half_jac = 0.5 * z.jac[p];
    const double sp = half_jac * (Uk + lambda0), sm = half_jac * (Uk - lambda0);
    const double pk = z.bv[p + 4 * sz] * half_jac;
    FP(0, ps, ex) = sp * q0;
    FP(1, ps, ex) = fma(sp, q1, pk * m0);
    FP(2, ps, ex) = fma(sp, q2, pk * m1);
    FP(3, ps, ex) = fma(sp, q3, pk * m2);
    FP(4, ps, ex) = fma(sp, q4, pk * Uk);
    FM(0, ps, ex) = sm * q0;
    FM(1, ps, ex) = fma(sm, q1, pk * m0);
    FM(2, ps, ex) = fma(sm, q2, pk * m1);
    FM(3, ps, ex) = fma(sm, q3, pk * m2);
    FM(4, ps, ex) = fma(sm, q4, pk * Uk);
#pragma unroll
    for (int l = 5; l < NV; ++l) {
      const double ql = z.cv[p + l * sz];
      FP(l, ps, ex) = sp * ql;
      FM(l, ps, ex) = sm * ql;
    }
  }
  __syncthreads();

  // ---- face fluxes ----
  const int face = f0 - 1 + ts;
  int fi, fj, fk;
  bool lat_ok;
  if constexpr (DIR == 0) {
    fi = face;
    fj = lj + tx;
    fk = lk;
    lat_ok = fj < z.my;
  } else if constexpr (DIR == 1) {
    fi = li + tx;
    fj = face;
    fk = lk;
    lat_ok = fi < z.mx;
  } else {
    fi = li + tx;
    fj = lj;
    fk = face;
    lat_ok = fi < z.mx;
  }
  const bool face_ok = lat_ok && face <= n_sweep - 1;
  double fc[NV];
  if (face_ok) {
    const long pf = z.ix(fi, fj, fk);
    const bool shock = face_shock(z, o, pf, stride);
    const double vr2 = o.v_ref * o.v_ref;
#pragma unroll 1
    for (int l = 0; l < NV; ++l) {
      const double eps_here = (l >= 1 && l <= 3) ? eps_ref * vr2 : (l == 4 ? eps_ref * vr2 * vr2 : eps_ref);
      double vp[2 * LO + 1], vm[2 * LO + 1];
#pragma unroll
      for (int m = 0; m < 2 * LO + 1; ++m) {
        vp[m] = FP(l, ts + m, tx);
        vm[m] = FM(l, ts + m + 1, tx);
      }
      const double f = weno_flux<W7>(vp, vm, eps_here, shock);
#pragma unroll
      for (int c = 0; c < NV; ++c)
        if (c == l) fc[c] = f;
    }
    if constexpr (NV > 5) {
      if (o.pp) {
        // the first-order flux's advection speeds take the xi metric (components
        // 0..2) in every direction, src/WENO.cu:1017-1045, 1746-1774, 2411-2439
        const long pl = pf, pr = pf + stride;
        const double jl = z.jac[pl], jr = z.jac[pr];
        const double t1 = (z.metric[pl] * z.cv[pl + sz] + z.metric[pl + sz] * z.cv[pl + 2 * sz] +
                           z.metric[pl + 2 * sz] * z.cv[pl + 3 * sz]) / z.cv[pl] * jl;
        const double t2 = (z.metric[pr] * z.cv[pr + sz] + z.metric[pr + sz] * z.cv[pr + 2 * sz] +
                           z.metric[pr + 2 * sz] * z.cv[pr + 3 * sz]) / z.cv[pr] * jr;
        const double dt = pp_time_step(z, o, fi, fj, fk);
#pragma unroll
        for (int l = 5; l < NV; ++l) {
          const double rhoYL = z.cv[pl + l * sz], rhoYR = z.cv[pr + l * sz];
          fc[l] = pp_limit(fc[l], 0.5 * o.alpha * jl, 0.5 * o.alpha * jr, rhoYL, rhoYR, dt,
                           FP(l, ts + LO, tx) + t1 * rhoYL, FM(l, ts + LO + 1, tx) - t2 * rhoYR, o.rho_ref);
        }
      }
    }
  }
  __syncthreads();  // every staged value is read
  if (face_ok) {
#pragma unroll
    for (int l = 0; l < NV; ++l) s[l * NT + me] = fc[l];
  }
  __syncthreads();
  if (face_ok && ts >= 1) {
    const long q = z.dx(fi, fj, fk);
    const long dsz = z.dsz;
    const int left = (ts - 1) * TX + tx;
#pragma unroll
    for (int l = 0; l < NV; ++l) z.dq[q + l * dsz] -= s[l * NT + me] - s[l * NT + left];
  }
#undef FP
#undef FM
}

// ---------------------------------------------------------------------------
// Shock sensors, src/FieldOperation.cu:122-205, written over the interior by
// compute_shock_sensor (207-228): 0 modified Ducros, 1 modified Jameson,
// 2 density and pressure jump.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_shock_sensor(Zone z, double *sensor, int kind, double eps, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz;
  const long st[3] = {1, z.sx, z.sxy};
  double out;
  if (kind == 1) {
    const double *pr = z.bv + 4 * sz;
    const double twoP = pr[p] * 2;
    const double pI = pr[p + 1] + pr[p - 1], pJ = pr[p + z.sx] + pr[p - z.sx], pK = pr[p + z.sxy] + pr[p - z.sxy];
    out = fabs(pI - twoP) / (pI + twoP) + fabs(pJ - twoP) / (pJ + twoP) + fabs(pK - twoP) / (pK + twoP);
  } else if (kind == 2) {
    const double *rho = z.bv, *pr = z.bv + 4 * sz;
    const double rhoI = 1.0 / rho[p], pI = 1.0 / pr[p];
    out = 0;
    out = fmax(out, fabs(rho[p + 1] - rho[p]) * rhoI);
    out = fmax(out, fabs(rho[p + z.sx] - rho[p]) * rhoI);
    out = fmax(out, fabs(rho[p + z.sxy] - rho[p]) * rhoI);
    out = fmax(out, fabs(pr[p + 1] - pr[p]) * pI);
    out = fmax(out, fabs(pr[p + z.sx] - pr[p]) * pI);
    out = fmax(out, fabs(pr[p + z.sxy] - pr[p]) * pI);
  } else {
    constexpr double a = 1.0 / 12.0, b = 2.0 / 3.0;
    double d[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const double *v = z.bv + (c + 1) * sz;
        d[c][e] = a * v[p - 2 * st[e]] - b * v[p - st[e]] + b * v[p + st[e]] - a * v[p + 2 * st[e]];
      }
    const double *m = z.metric;
    const double xi_x = m[p], xi_y = m[p + sz], xi_z = m[p + 2 * sz];
    const double eta_x = m[p + 3 * sz], eta_y = m[p + 4 * sz], eta_z = m[p + 5 * sz];
    const double zeta_x = m[p + 6 * sz], zeta_y = m[p + 7 * sz], zeta_z = m[p + 8 * sz];
    const double duDx = d[0][0] * xi_x + d[0][1] * eta_x + d[0][2] * zeta_x;
    const double dvDy = d[1][0] * xi_y + d[1][1] * eta_y + d[1][2] * zeta_y;
    const double dwDz = d[2][0] * xi_z + d[2][1] * eta_z + d[2][2] * zeta_z;
    double div2 = duDx + dvDy + dwDz;
    if (div2 > 0) {
      out = 0;
    } else {
      const double duDy = d[0][0] * xi_y + d[0][1] * eta_y + d[0][2] * zeta_y;
      const double duDz = d[0][0] * xi_z + d[0][1] * eta_z + d[0][2] * zeta_z;
      const double dvDx = d[1][0] * xi_x + d[1][1] * eta_x + d[1][2] * zeta_x;
      const double dvDz = d[1][0] * xi_z + d[1][1] * eta_z + d[1][2] * zeta_z;
      const double dwDx = d[2][0] * xi_x + d[2][1] * eta_x + d[2][2] * zeta_x;
      const double dwDy = d[2][0] * xi_y + d[2][1] * eta_y + d[2][2] * zeta_y;
      div2 *= div2;
      const double curl2 =
          (dwDy - dvDz) * (dwDy - dvDz) + (duDz - dwDx) * (duDz - dwDx) + (dvDx - duDy) * (dvDx - duDy);
      out = div2 / (div2 + curl2 + eps);
    }
  }
  sensor[p] = out;
}

// ---------------------------------------------------------------------------
// Primitive-variable inviscid path: reconstruction (first order, MUSCL, NND2
// with minmod), then the Lax-Friedrichs (Air), AUSM+ or HLLC solver, or the
// Roe solver with its entropy fix; src/InviscidScheme.cu:61-219, 269-395,
// 397-787, src/Reconstruction.cu.  Same block decomposition as the WENO
// kernels: TS faces per line, one face overlap between blocks, face fluxes
// exchanged through the staging buffer for the difference.
// ---------------------------------------------------------------------------
struct PvArgs {
  int solver;  // 1 LF, 2 Roe, 3 AUSM+, 4 HLLC
  int recon;   // 1 first order, 2 MUSCL, 3 NND2, 4 WENO5-JS, 5 WENO7-Z (primitive variables)
  double efix; // Roe entropy-fix factor
  int dim;     // 2 or 3
  double eps_rho, eps_vel, eps_p;  // WENO epsilon of rho, u/v/w, p (1e-40 times their squared scales)
};

DEV double sgn1(double a) { return a < 0 ? -1.0 : 1.0; }  // gxl::sgn
// Limiter<0/1>::apply_limiter, src/Limiter.cuh:14-16 (minmod for both)
DEV double minmod(double a, double b) { return 0.5 * (sgn1(a) + sgn1(b)) * fmin(fabs(a), fabs(b)); }

// e_tr and gamma of the mixture at t (compute_mixture_characteristic_thermo, src/Thermo.cuh:49-87)
template <int M>
DEV void mix_e_gamma(const Tab *__restrict__ tb, double t, const double *y, double &e_tr, double &gamma,
                     double &c) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const double t_eval = fmax(t, 1.0);
  double h[NS], cp[NS];
  thermo_h_cp<M>(tb, t_eval, h, cp);
  if constexpr (kTT(M)) {
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      double cv_ve;
      h[l] -= ve_energy1<true>(tb, l, t_eval, &cv_ve);
      cp[l] -= cv_ve;
    }
  }
  double e = 0, rl = 0, cpl = 0, cvl = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    e += y[l] * (h[l] - tb->gas_const[l] * t_eval);
    rl += y[l] * tb->gas_const[l];
    cpl += y[l] * cp[l];
    cvl += y[l] * (cp[l] - tb->gas_const[l]);
  }
  gamma = cpl / fmax(cvl, 1e-8);
  e_tr = e;
  c = sqrt(fmax(gamma * rl * t_eval, 1e-12));
}

// RW: nodes staged on each side of a face: 2 for the first-order / MUSCL /
// NND2 window i-1 .. i+2, 3 for WENO5 (i-2 .. i+3), 4 for WENO7 (i-3 .. i+4).
template <int M, int DIR, int TX, int TS, int RW>
__global__ __launch_bounds__(TX *TS) void k_pv(Zone z, const Tab *__restrict__ tb, PvArgs a) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  constexpr int NSC = kNsc(M);  // species (and E_ve)
  constexpr int NV = kNv(M);
  constexpr int NP = TS + 2 * RW - 1;  // points of the TS faces: f0-RW .. f0+TS+RW-2
  constexpr int NQ = NV + 5;   // rho u v w p Y | metric[3] jac entropy-fix delta
  constexpr int NT = TX * TS;
  constexpr int NSTAGE = NQ * NP * TX, NFC = NV * NT;
  __shared__ double s[NSTAGE > NFC ? NSTAGE : NFC];
#define SP(v, ps, tx) s[((v) * NP + (ps)) * TX + (tx)]
  const int tid = threadIdx.x;
  int tx, ts;
  if constexpr (DIR == 0) {
    ts = tid % TS;
    tx = tid / TS;
  } else {
    tx = tid % TX;
    ts = tid / TX;
  }
  const int me = ts * TX + tx;
  const int n_sweep = DIR == 0 ? z.mx : (DIR == 1 ? z.my : z.mz);
  int f0, li, lj, lk;
  int bx, by, bz;
  sweep_tile<DIR>(bx, by, bz);
  if constexpr (DIR == 0) {
    f0 = bx * (TS - 1);
    lj = by * TX;
    lk = bz;
    li = 0;
  } else if constexpr (DIR == 1) {
    li = bx * TX;
    f0 = by * (TS - 1);
    lk = bz;
    lj = 0;
  } else {
    li = bx * TX;
    lj = by;
    f0 = bz * (TS - 1);
    lk = 0;
  }
  const long sz = z.sz;
  for (int e = tid; e < NP * TX; e += NT) {
    int ps, ex;
    if constexpr (DIR == 0) {
      ps = e % NP;
      ex = e / NP;
    } else {
      ex = e % TX;
      ps = e / TX;
    }
    const int pos = f0 - RW + ps;
    int i, j, k;
    bool lat_ok;
    if constexpr (DIR == 0) {
      i = pos;
      j = lj + ex;
      k = lk;
      lat_ok = j < z.my;
    } else if constexpr (DIR == 1) {
      i = li + ex;
      j = pos;
      k = lk;
      lat_ok = i < z.mx;
    } else {
      i = li + ex;
      j = lj;
      k = pos;
      lat_ok = i < z.mx;
    }
    if (!lat_ok || pos > n_sweep - 1 + RW) continue;
    const long p = z.ix(i, j, k);
#pragma unroll
    for (int l = 0; l < 5; ++l) SP(l, ps, ex) = z.bv[p + l * sz];
#pragma unroll
    for (int l = 0; l < NSC; ++l) SP(5 + l, ps, ex) = z.sv[p + l * sz];
#pragma unroll
    for (int l = 0; l < 3; ++l) SP(NV + l, ps, ex) = z.metric[p + (3 * DIR + l) * sz];
    SP(NV + 3, ps, ex) = z.jac[p];
    double efd = 0;
    if (a.solver == 2) {
      // compute_entropy_fix_delta, src/InviscidScheme.cu:269-304
      const double *mt = z.metric;
      const double u = z.bv[p + sz], v = z.bv[p + 2 * sz], w = z.bv[p + 3 * sz];
      const double U = fabs(u * mt[p] + v * mt[p + sz] + w * mt[p + 2 * sz]);
      const double V = fabs(u * mt[p + 3 * sz] + v * mt[p + 4 * sz] + w * mt[p + 5 * sz]);
      const double W = fabs(u * mt[p + 6 * sz] + v * mt[p + 7 * sz] + w * mt[p + 8 * sz]);
      const double kx = norm3(mt[p], mt[p + sz], mt[p + 2 * sz]);
      const double ky = norm3(mt[p + 3 * sz], mt[p + 4 * sz], mt[p + 5 * sz]);
      const double kz = norm3(mt[p + 6 * sz], mt[p + 7 * sz], mt[p + 8 * sz]);
      const double ac = NS > 0 ? z.acoustic[p] : sqrt(kGammaAir * z.bv[p + 4 * sz] / z.bv[p]);
      efd = a.dim == 2 ? a.efix * (U + V + ac * 0.5 * (kx + ky)) : a.efix * (U + V + W + ac * (kx + ky + kz) / 3.0);
    }
    SP(NV + 4, ps, ex) = efd;
  }
  __syncthreads();

  const int face = f0 - 1 + ts;
  int fi, fj, fk;
  bool lat_ok;
  if constexpr (DIR == 0) {
    fi = face;
    fj = lj + tx;
    fk = lk;
    lat_ok = fj < z.my;
  } else if constexpr (DIR == 1) {
    fi = li + tx;
    fj = face;
    fk = lk;
    lat_ok = fi < z.mx;
  } else {
    fi = li + tx;
    fj = lj;
    fk = face;
    lat_ok = fi < z.mx;
  }
  const bool face_ok = lat_ok && face <= n_sweep - 1;
  double fc[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) fc[l] = 0;
  if (face_ok) {
    const int b1 = ts + RW - 1, b2 = b1 + 1;  // the face lies between window points b1 and b2
    // ---- reconstruction<mix>, src/InviscidScheme.cu:171-219, src/Reconstruction.cu:5-44 ----
    double pl[NV + 2], pr[NV + 2];
    bool weno_ok = true;
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      const double q0 = SP(l, b1 - 1, tx), q1 = SP(l, b1, tx), q2 = SP(l, b2, tx), q3 = SP(l, b2 + 1, tx);
      if constexpr (RW >= 3) {
        // WENO5-JS (RW 3) / WENO7-Z (RW 4) of the primitive variables (no reference
        // counterpart: the pairing of a WENO reconstruction with the AUSM+ / HLLC / Roe solvers)
        const double eps = l == 0 ? a.eps_rho : (l < 4 ? a.eps_vel : (l == 4 ? a.eps_p : 1e-40));
        double v[2 * RW - 1], w[2 * RW - 1];
        if constexpr (RW == 4) {
#pragma unroll
          for (int m = 0; m < 7; ++m) {
            v[m] = SP(l, b1 - 3 + m, tx);
            w[m] = SP(l, b2 + 3 - m, tx);
          }
          pl[l] = weno7_half(v[0], v[1], v[2], v[3], v[4], v[5], v[6], eps);
          pr[l] = weno7_half(w[0], w[1], w[2], w[3], w[4], w[5], w[6], eps);
        } else {
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            v[m] = SP(l, b1 - 2 + m, tx);
            w[m] = SP(l, b1 - 1 + m, tx);
          }
          pl[l] = weno5_left(v, eps);
          pr[l] = weno5_right(w, eps);
        }
        if (l == 0 || l == 4) weno_ok = weno_ok && pl[l] > 0 && pr[l] > 0;
      } else if (a.recon == 2) {
        constexpr double kappa = 1.0 / 3.0;
        const double di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
        const double neg_l = minmod(di, di1), pos_l = minmod(di1, di);
        const double neg_r = minmod(di1, di2), pos_r = minmod(di2, di1);
        pl[l] = q1 + 0.25 * ((1 - kappa) * neg_l + (1 + kappa) * pos_l);
        pr[l] = q2 - 0.25 * ((1 - kappa) * pos_r + (1 + kappa) * neg_r);
      } else if (a.recon == 3) {
        const double di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
        pl[l] = q1 + 0.5 * minmod(di, di1);
        pr[l] = q2 - 0.5 * minmod(di1, di2);
      } else {
        pl[l] = q1;
        pr[l] = q2;
      }
    }
    if constexpr (RW >= 3) {
      // a WENO state with a non-positive density or pressure falls back to first order
      if (!weno_ok) {
#pragma unroll
        for (int l = 0; l < NV; ++l) {
          pl[l] = SP(l, b1, tx);
          pr[l] = SP(l, b2, tx);
        }
      }
    }
    double gam_l = kGammaAir, gam_r = kGammaAir;
    if constexpr (NS > 0) {
      const double ek_l = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
      const double ek_r = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
      double rl = 0, rr = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        rl += pl[5 + l] * tb->gas_const[l];
        rr += pr[5 + l] * tb->gas_const[l];
      }
      double e_l, e_r, c_unused;
      mix_e_gamma<M>(tb, pl[4] / fmax(pl[0] * rl, 1e-12), pl + 5, e_l, gam_l, c_unused);
      mix_e_gamma<M>(tb, pr[4] / fmax(pr[0] * rr, 1e-12), pr + 5, e_r, gam_r, c_unused);
      double eve_l = 0.0, eve_r = 0.0;
      if constexpr (kTT(M)) {
        eve_l = fmax(pl[5 + NS], 0.0);
        eve_r = fmax(pr[5 + NS], 0.0);
      }
      pl[NV] = pl[0] * (ek_l + e_l + eve_l);
      pr[NV] = pr[0] * (ek_r + e_r + eve_r);
    } else {
      pl[NV] = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) * pl[0] + pl[4] / (kGammaAir - 1);
      pr[NV] = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) * pr[0] + pr[4] / (kGammaAir - 1);
    }
    const double mlx = SP(NV, b1, tx), mly = SP(NV + 1, b1, tx), mlz = SP(NV + 2, b1, tx);
    const double mrx = SP(NV, b2, tx), mry = SP(NV + 1, b2, tx), mrz = SP(NV + 2, b2, tx);
    const double jl = SP(NV + 3, b1, tx), jr = SP(NV + 3, b2, tx);
    if (a.solver == 3) {
      // riemannSolver_ausmPlus, src/InviscidScheme.cu:397-469
      const double k1 = 0.5 * (jl * mlx + jr * mrx), k2 = 0.5 * (jl * mly + jr * mry), k3 = 0.5 * (jl * mlz + jr * mrz);
      const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
      const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk, ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
      const double c = 0.5 * (sqrt(gam_l * pl[4] / pl[0]) + sqrt(gam_r * pr[4] / pr[0]));
      const double ml = ul / c, mr = ur / c;
      constexpr double alpha = 3.0 / 16.0;
      double mlp, plp, mrn, prn;
      if (fabs(ml) > 1) {
        mlp = 0.5 * (ml + fabs(ml));
        plp = mlp / ml;
      } else {
        const double q = (ml + 1) * (ml + 1) * 0.25, b = (ml * ml - 1) * (ml * ml - 1);
        mlp = q + 0.125 * b;
        plp = q * (2 - ml) + alpha * ml * b;
      }
      if (fabs(mr) > 1) {
        mrn = 0.5 * (mr - fabs(mr));
        prn = mrn / mr;
      } else {
        const double q = (mr - 1) * (mr - 1) * 0.25, b = (mr * mr - 1) * (mr * mr - 1);
        mrn = -q - 0.125 * b;
        prn = q * (2 + mr) - alpha * mr * b;
      }
      const double p_coeff = plp * pl[4] + prn * pr[4];
      const double m_half = mlp + mrn;
      const double mass = c * (pl[0] * (0.5 * (m_half + fabs(m_half))) + pr[0] * (0.5 * (m_half - fabs(m_half))));
      const double coeff = mass * gk;
      const double *q = mass >= 0 ? pl : pr;
      fc[0] = coeff;
      fc[1] = coeff * q[1] + p_coeff * k1;
      fc[2] = coeff * q[2] + p_coeff * k2;
      fc[3] = coeff * q[3] + p_coeff * k3;
      fc[4] = coeff * (q[NV] + q[4]) / q[0];
#pragma unroll
      for (int l = 5; l < NV; ++l) fc[l] = coeff * q[l];
    } else if (a.solver == 1) {
      // riemannSolver_laxFriedrich<Air>, src/InviscidScheme.cu:753-787
      const double kx = 0.5 * (mlx + mrx), ky = 0.5 * (mly + mry), kz = 0.5 * (mlz + mrz);
      const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
      const double Ukl = pl[1] * kx + pl[2] * ky + pl[3] * kz, Ukr = pr[1] * kx + pr[2] * ky + pr[3] * kz;
      const double cl = sqrt(kGammaAir * pl[4] / pl[0]), cr = sqrt(kGammaAir * pr[4] / pr[0]);
      const double sr = fmax(fabs(Ukl) + cl * gradK, fabs(Ukr) + cr * gradK);
      const double hj = 0.5 * 0.5 * (jl + jr);
      const double rul = pl[0] * Ukl, rur = pr[0] * Ukr;
      fc[0] = (rul + rur - sr * (pr[0] - pl[0])) * hj;
      fc[1] = (rul * pl[1] + rur * pr[1] + kx * (pl[4] + pr[4]) - sr * (pr[1] * pr[0] - pl[1] * pl[0])) * hj;
      fc[2] = (rul * pl[2] + rur * pr[2] + ky * (pl[4] + pr[4]) - sr * (pr[2] * pr[0] - pl[2] * pl[0])) * hj;
      fc[3] = (rul * pl[3] + rur * pr[3] + kz * (pl[4] + pr[4]) - sr * (pr[3] * pr[0] - pl[3] * pl[0])) * hj;
      fc[4] = ((pl[NV] + pl[4]) * Ukl + (pr[NV] + pr[4]) * Ukr - sr * (pr[NV] - pl[NV])) * hj;
    } else if (a.solver == 4) {
      // riemannSolver_hllc, src/InviscidScheme.cu:471-590
      const double sl0 = sqrt(pl[0]), sr0 = sqrt(pr[0]);
      const double rl_c = sl0 / (sl0 + sr0), rr_c = sr0 / (sl0 + sr0);
      const double ut = pl[1] * rl_c + pr[1] * rr_c, vt = pl[2] * rl_c + pr[2] * rr_c, wt = pl[3] * rl_c + pr[3] * rr_c;
      double c_t;
      if constexpr (NS == 0) {
        const double ht = (pl[NV] + pl[4]) / pl[0] * rl_c + (pr[NV] + pr[4]) / pr[0] * rr_c;
        c_t = sqrt((kGammaAir - 1) * (ht - 0.5 * (ut * ut + vt * vt + wt * wt)));
      } else {
        double svm[NS];
        double rg = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          svm[l] = rl_c * pl[l + 5] + rr_c * pr[l + 5];
          rg += svm[l] * tb->gas_const[l];
        }
        const double t = (rl_c * (pl[4] / pl[0]) + rr_c * (pr[4] / pr[0])) / fmax(rg, 1e-12);
        double e_u, g_u;
        mix_e_gamma<M>(tb, t, svm, e_u, g_u, c_t);
      }
      const double kx = 0.5 * (mlx + mrx), ky = 0.5 * (mly + mry), kz = 0.5 * (mlz + mrz);
      const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
      const double Ut = (kx * ut + ky * vt + kz * wt) / gradK;
      const double jac_ave = 0.5 * (jl + jr);
      const double Ul = kx * pl[1] + ky * pl[2] + kz * pl[3], Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
      const double sl = fmin(Ul / gradK - sqrt(gam_l * pl[4] / pl[0]), Ut - c_t);
      const double sr = fmax(Ur / gradK + sqrt(gam_r * pr[4] / pr[0]), Ut + c_t);
      if (sl >= 0 || sr < 0) {
        const double *q = sl >= 0 ? pl : pr;
        const double U = sl >= 0 ? Ul : Ur;
        const double ru = q[0] * U;
        fc[0] = jac_ave * ru;
        fc[1] = jac_ave * (ru * q[1] + q[4] * kx);
        fc[2] = jac_ave * (ru * q[2] + q[4] * ky);
        fc[3] = jac_ave * (ru * q[3] + q[4] * kz);
        fc[4] = jac_ave * ((q[NV] + q[4]) * U);
#pragma unroll
        for (int l = 5; l < NV; ++l) fc[l] = jac_ave * ru * q[l];
      } else {
        const double sm = ((pr[0] * Ur * (sr - Ur / gradK) - pl[0] * Ul * (sl - Ul / gradK)) / gradK + pl[4] - pr[4]) /
                          (pr[0] * (sr - Ur / gradK) - pl[0] * (sl - Ul / gradK));
        const double pm = pl[0] * (sl - Ul / gradK) * (sm - Ul / gradK) + pl[4];
        const double *q = sm >= 0 ? pl : pr;
        const double sw = sm >= 0 ? sl : sr, U = sm >= 0 ? Ul : Ur;
        const double pc = 1.0 / (sw - sm);
        const double Qc = jac_ave * pc * sm * (sw * gradK - U) * q[0];
        const double dP = (sw * pm - sm * q[4]) * pc * jac_ave;
        fc[0] = Qc;
        fc[1] = Qc * q[1] + dP * kx;
        fc[2] = Qc * q[2] + dP * ky;
        fc[3] = Qc * q[3] + dP * kz;
        fc[4] = Qc * q[NV] / q[0] + pc * jac_ave * (sw * pm * sm * gradK - sm * q[4] * U);
#pragma unroll
        for (int l = 5; l < NV; ++l) fc[l] = Qc * q[l];
      }
    } else {
      // riemannSolver_Roe with compute_half_sum_left_right_flux, src/InviscidScheme.cu:592-746
      {
        const double jx = jl * mlx, jy = jl * mly, jz = jl * mlz;
        const double Uk = pl[1] * jx + pl[2] * jy + pl[3] * jz, coeff = Uk * pl[0];
        fc[0] = 0.5 * coeff;
        fc[1] = 0.5 * (coeff * pl[1] + pl[4] * jx);
        fc[2] = 0.5 * (coeff * pl[2] + pl[4] * jy);
        fc[3] = 0.5 * (coeff * pl[3] + pl[4] * jz);
        fc[4] = 0.5 * Uk * (pl[4] + pl[NV]);
#pragma unroll
        for (int l = 5; l < NV; ++l) fc[l] = 0.5 * coeff * pl[l];
      }
      {
        const double jx = jr * mrx, jy = jr * mry, jz = jr * mrz;
        const double Uk = pr[1] * jx + pr[2] * jy + pr[3] * jz, coeff = Uk * pr[0];
        fc[0] += 0.5 * coeff;
        fc[1] += 0.5 * (coeff * pr[1] + pr[4] * jx);
        fc[2] += 0.5 * (coeff * pr[2] + pr[4] * jy);
        fc[3] += 0.5 * (coeff * pr[3] + pr[4] * jz);
        fc[4] += 0.5 * Uk * (pr[4] + pr[NV]);
#pragma unroll
        for (int l = 5; l < NV; ++l) fc[l] += 0.5 * coeff * pr[l];
      }
      const double dl = sqrt(pl[0]), dr = sqrt(pr[0]);
      const double inv = 1.0 / (dl + dr);
      const double u = (dl * pl[1] + dr * pr[1]) * inv, v = (dl * pl[2] + dr * pr[2]) * inv,
                   w = (dl * pl[3] + dr * pr[3]) * inv;
      const double ek = 0.5 * (u * u + v * v + w * w);
      const double h = (dl * ((pl[NV] + pl[4]) / pl[0]) + dr * ((pr[NV] + pr[4]) / pr[0])) * inv;
      double gamma = kGammaAir;
      double c = sqrt((gamma - 1) * (h - ek));
      double svm[NSC > 0 ? NSC : 1];
      double t_roe = 1, c2s = 1, inv_gm1 = 0;
      if constexpr (NS > 0) {
        double rg = 0;
#pragma unroll
        for (int l = 0; l < NSC; ++l) svm[l] = (dl * pl[l + 5] + dr * pr[l + 5]) * inv;
#pragma unroll
        for (int l = 0; l < NS; ++l) rg += svm[l] * tb->gas_const[l];
        t_roe = fmax((dl * (pl[4] / pl[0]) + dr * (pr[4] / pr[0])) * inv / fmax(rg, 1e-12), 1.0);
        double e_u;
        mix_e_gamma<M>(tb, t_roe, svm, e_u, gamma, c);
        c2s = fmax(c * c, 1e-30);
        inv_gm1 = 1.0 / fmax(gamma - 1.0, 1e-12);
      }
      double kx = 0.5 * (mlx + mrx), ky = 0.5 * (mly + mry), kz = 0.5 * (mlz + mrz);
      const double gradK = sqrt(kx * kx + ky * ky + kz * kz);
      double Uk = kx * u + ky * v + kz * w;
      double ch0 = fabs(Uk - gradK * c), ch1 = fabs(Uk), ch2 = fabs(Uk + gradK * c);
      const double efd = 0.5 * (SP(NV + 4, b1, tx) + SP(NV + 4, b2, tx));
      if (ch0 < efd) ch0 = 0.5 * (ch0 * ch0 / efd + efd);
      if (ch1 < efd) ch1 = 0.5 * (ch1 * ch1 / efd + efd);
      if (ch2 < efd) ch2 = 0.5 * (ch2 * ch2 / efd + efd);
      kx /= gradK;
      ky /= gradK;
      kz /= gradK;
      Uk /= gradK;
      const double jac_ave = 0.5 * (jl + jr);
      double dq[NV];
      dq[0] = jac_ave * (pr[0] - pl[0]);
#pragma unroll
      for (int l = 1; l < NV; ++l) dq[l] = jac_ave * (pr[0] * pr[l] - pl[0] * pl[l]);
      dq[4] = jac_ave * (pr[NV] - pl[NV]);
      double c1 = (gamma - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
      double c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - Uk * dq[0]) / c;
      double h_s[NS > 0 ? NS : 1];
      if constexpr (NS > 0) {
        const double te2 = t_roe * t_roe, te3 = te2 * t_roe, te4 = te3 * t_roe, te5 = te4 * t_roe;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          h_s[l] = thermo_h1<M>(tb, l, t_roe, te2, te3, te4, te5);
          if constexpr (kTT(M)) h_s[l] -= ve_energy1<false>(tb, l, t_roe, nullptr);
          const double alpha = (gamma * tb->gas_const[l] * t_roe - (gamma - 1.0) * h_s[l]) / c2s;
          c1 += alpha * dq[l + 5];
        }
        if constexpr (kTT(M)) c1 += -(gamma - 1.0) / c2s * dq[5 + NS];
      }
      double c3 = dq[0] - c1;
      double b[NV];
      const double L0 = 0.5 * (c1 - c2), L4 = 0.5 * (c1 + c2);
      b[0] = -ch0 * L0;
      b[1] = -ch1 * (kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c);
      b[2] = -ch1 * (ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c);
      b[3] = -ch1 * (kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c);
      b[4] = -ch2 * L4;
#pragma unroll
      for (int l = 0; l < NSC; ++l) b[l + 5] = -ch1 * (dq[l + 5] - svm[l] * dq[0]);
      const double c0 = kx * b[1] + ky * b[2] + kz * b[3];
      c1 = c0 + b[0] + b[4];
      c2 = c * (b[4] - b[0]);
      c3 = 0;
      if constexpr (NS > 0) {
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const double alpha = (gamma * tb->gas_const[l] * t_roe - (gamma - 1.0) * h_s[l]) / c2s;
          c3 += -alpha * c2s * inv_gm1 * b[l + 5];
        }
        if constexpr (kTT(M)) c3 += 1.0 * b[5 + NS];
      }
      fc[0] += 0.5 * c1;
      fc[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * b[2] - ky * b[3]));
      fc[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * b[3] - kz * b[1]));
      fc[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * b[1] - kx * b[2]));
      fc[4] += 0.5 * (h * c1 + Uk * c2 - c * c * c0 / (gamma - 1) +
                      c * ((kz * v - ky * w) * b[1] + (kx * w - kz * u) * b[2] + (ky * u - kx * v) * b[3]) + c3);
#pragma unroll
      for (int l = 0; l < NSC; ++l) fc[5 + l] += 0.5 * (b[l + 5] + svm[l] * c1);
    }
  }
  __syncthreads();
#pragma unroll
  for (int l = 0; l < NV; ++l) s[l * NT + me] = fc[l];
  __syncthreads();
  if (face_ok && ts >= 1) {
    const long q = z.dx(fi, fj, fk);
    const int left = (ts - 1) * TX + tx;
#pragma unroll
    for (int l = 0; l < NV; ++l) z.dq[q + l * z.dsz] -= s[l * NT + me] - s[l * NT + left];
  }
#undef SP
}

// ---------------------------------------------------------------------------
// 8th-order collocated viscous fluxes, src/ViscousScheme.cu:140-834
// ---------------------------------------------------------------------------
// 8th-order first derivative at node i of n along one axis, with the
// reference's one-sided closures next to a physical boundary (d_dXi,
// src/ViscousScheme.cu:140-188).  F(o) is the value o nodes away.
template <class F>
DEV double d8g(F F_, int i, int n, bool bl, bool br) {
  if (bl && i <= 3) {
    if (!(br && i >= n - 4)) {
      if (i == 0) return -25.0 / 12 * F_(0) + 4.0 * F_(1) - 3 * F_(2) + 4.0 / 3 * F_(3) - 0.25 * F_(4);
      if (i == 1) return -0.25 * F_(-1) - 5.0 / 6 * F_(0) + 1.5 * F_(1) - 0.5 * F_(2) + 1.0 / 12 * F_(3);
      if (i == 2) return -1.0 / 12 * (F_(2) - F_(-2)) + 2.0 / 3 * (F_(1) - F_(-1));
      return 0.75 * (F_(1) - F_(-1)) - 0.15 * (F_(2) - F_(-2)) + 1.0 / 60 * (F_(3) - F_(-3));
    }
  }
  if (br && i >= n - 4) {
    if (i == n - 1) return 25.0 / 12 * F_(0) - 4.0 * F_(-1) + 3 * F_(-2) - 4.0 / 3 * F_(-3) + 0.25 * F_(-4);
    if (i == n - 2) return 0.25 * F_(1) + 5.0 / 6 * F_(0) - 1.5 * F_(-1) + 0.5 * F_(-2) - 1.0 / 12 * F_(-3);
    if (i == n - 3) return -1.0 / 12 * (F_(2) - F_(-2)) + 2.0 / 3 * (F_(1) - F_(-1));
    return 0.75 * (F_(1) - F_(-1)) - 0.15 * (F_(2) - F_(-2)) + 1.0 / 60 * (F_(3) - F_(-3));
  }
  constexpr double a1 = 0.8, a2 = -0.2, a3 = 4.0 / 105, a4 = -1.0 / 280;
  return a1 * (F_(1) - F_(-1)) + a2 * (F_(2) - F_(-2)) + a3 * (F_(3) - F_(-3)) + a4 * (F_(4) - F_(-4));
}

DEV double d8(const double *__restrict__ f, long p, long s, int i, int n, bool bl, bool br) {
  return d8g([&](int o) { return f[p + static_cast<long>(o) * s]; }, i, n, bl, br);
}


DEV void btype_flags(const Zone &z, int i, int j, int k, bool *ct) {
  ct[0] = z.bt_il[j + k * z.my] != 0;
  ct[1] = z.bt_ir[j + k * z.my] != 0;
  ct[2] = z.bt_jl[i + k * z.mx] != 0;
  ct[3] = z.bt_jr[i + k * z.mx] != 0;
  ct[4] = z.bt_kl[i + j * z.mx] != 0;
  ct[5] = z.bt_kr[i + j * z.mx] != 0;
}

// The zeta derivatives (8th order with the boundary closures, the DZETA
// terms of src/ViscousScheme.cu:470-788) of the fields the viscous fluxes
// differentiate, u, v, w, T, Y_l, with the pressure-gradient term p and with
// two temperatures Tve (field tve_f), into
// zg[field] at interior nodes.  A thread owns KC consecutive k of one column
// and slides a nine-node window along them, so each field is read
// (KC + 8) / KC times instead of nine by the flux kernel.
template <int KC>
__global__ __launch_bounds__(kBlock) void k_zeta_grad(Zone z, int ns, int nfield, int tve_f, long ncol) {
  const long t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= ncol) return;
  const int mx = z.mx, my = z.my, mz = z.mz;
  const int i = static_cast<int>(t % mx);
  const long r = t / mx;
  const int j = static_cast<int>(r % my);
  const int k0 = static_cast<int>(r / my) * KC;
  const int nk = min(KC, mz - k0);
  const long sz = z.sz, sk = z.sxy, dsz = z.dsz, dk = static_cast<long>(mx) * my;
  const bool kl = z.bt_kl[i + j * mx] != 0, kr = z.bt_kr[i + j * mx] != 0;
  const long p0 = z.ix(i, j, k0), q0 = z.dx(i, j, k0);
#pragma unroll 1
  for (int f = 0; f < nfield; ++f) {
    const double *src = f == tve_f ? z.tve
                                   : (f < 3 ? z.bv + (f + 1) * sz
                                            : (f == 3 ? z.bv + 5 * sz : (f < 4 + ns ? z.sv + (f - 4) * sz : z.bv + 4 * sz)));
    // nodes k0-4 .. k0+nk+3 lie inside the padded extent (ngg >= 4)
    double w[9];
#pragma unroll
    for (int m = 0; m < 8; ++m) w[m] = src[p0 + (m - 4) * sk];
#pragma unroll 1
    for (int c = 0; c < nk; ++c) {
      w[8] = src[p0 + (c + 4) * sk];
      z.zg[q0 + c * dk + f * dsz] = d8g([&](int o) { return w[4 + o]; }, k0 + c, mz, kl, kr);
#pragma unroll
      for (int m = 0; m < 8; ++m) w[m] = w[m + 1];
    }
  }
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_viscous_flux(Zone z, const Tab *__restrict__ tb, double prandtl,
                                                         int gradp, int tve_f, long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  bool ct[6];
  btype_flags(z, i, j, k, ct);
  const long p = z.ix(i, j, k), sz = z.sz, si = 1, sj = z.sx, sk = z.sxy;
  const int mx = z.mx, my = z.my, mz = z.mz;
  const double *pv = z.bv;
#define DXI(arr, l) d8(arr, p + static_cast<long>(l) * sz, si, i, mx, ct[0], ct[1])
#define DETA(arr, l) d8(arr, p + static_cast<long>(l) * sz, sj, j, my, ct[2], ct[3])
#define DZETA(arr, l) d8(arr, p + static_cast<long>(l) * sz, sk, k, mz, ct[4], ct[5])
  const long dsz = z.dsz;
  const double u_xi = DXI(pv, 1), u_eta = DETA(pv, 1), u_zeta = z.zg[t];
  const double v_xi = DXI(pv, 2), v_eta = DETA(pv, 2), v_zeta = z.zg[t + dsz];
  const double w_xi = DXI(pv, 3), w_eta = DETA(pv, 3), w_zeta = z.zg[t + 2 * dsz];
  const double t_xi = DXI(pv, 5), t_eta = DETA(pv, 5), t_zeta = z.zg[t + 3 * dsz];
  const double *mt = z.metric;
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
  const double mul = z.mul[p];
  const double tau_xx = mul * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  const double tau_yy = mul * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  const double tau_zz = mul * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const double tau_xy = mul * (u_y + v_x);
  const double tau_xz = mul * (u_z + w_x);
  const double tau_yz = mul * (v_z + w_y);
  double *fv = z.ff, *gv = z.gf, *hv = z.hf;
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
  if constexpr (NS > 0) {
    conductivity = z.kappa[p];
  } else {
    constexpr double cp_air = kGammaAir * kRair / (kGammaAir - 1);
    conductivity = mul / prandtl * cp_air;
  }
  const double u = pv[p + sz], v = pv[p + 2 * sz], w = pv[p + 3 * sz];
  const double Ex = u * tau_xx + v * tau_xy + w * tau_xz + conductivity * t_x;
  const double Ey = u * tau_xy + v * tau_yy + w * tau_yz + conductivity * t_y;
  const double Ez = u * tau_xz + v * tau_yz + w * tau_zz + conductivity * t_z;
  double f4 = xi_x * Ex + xi_y * Ey + xi_z * Ez;
  double g4 = eta_x * Ex + eta_y * Ey + eta_z * Ez;
  double h4 = zeta_x * Ex + zeta_y * Ey + zeta_z * Ez;
  // two temperatures: kappa_ve grad Tve joins the energy flux and starts the
  // E_ve flux, src/ViscousScheme.cu:570-584, 703-719
  [[maybe_unused]] double eve_fx = 0, eve_fy = 0, eve_fz = 0;
  if constexpr (kTT(M)) {
    const double tve_xi = DXI(z.tve, 0), tve_eta = DETA(z.tve, 0), tve_zeta = z.zg[t + tve_f * dsz];
    const double tve_x = tve_xi * xi_x + tve_eta * eta_x + tve_zeta * zeta_x;
    const double tve_y = tve_xi * xi_y + tve_eta * eta_y + tve_zeta * zeta_y;
    const double tve_z = tve_xi * xi_z + tve_eta * eta_z + tve_zeta * zeta_z;
    const double kve = z.kappa_ve[p];
    eve_fx = kve * (xi_x * tve_x + xi_y * tve_y + xi_z * tve_z);
    eve_fy = kve * (eta_x * tve_x + eta_y * tve_y + eta_z * tve_z);
    eve_fz = kve * (zeta_x * tve_x + zeta_y * tve_y + zeta_z * tve_z);
    f4 += eve_fx;
    g4 += eve_fy;
    h4 += eve_fz;
  }
  if constexpr (NS > 0) {
    // species diffusion with correction velocity, src/ViscousScheme.cu:587-788
    const double *yv = z.sv;
    double diffusivity[NS], yk[NS], dfx[NS], dfy[NS], dfz[NS];
    double sGx = 0, sGy = 0, sGz = 0, sum_rhoDkYk = 0, mw_tot = 0;
    double CVx = 0, CVy = 0, CVz = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      yk[l] = yv[p + l * sz];
      diffusivity[l] = z.rhoD[p + l * sz];
      const double y_xi = DXI(yv, l), y_eta = DETA(yv, l), y_zeta = z.zg[t + (4 + l) * dsz];
      const double y_x = y_xi * xi_x + y_eta * eta_x + y_zeta * zeta_x;
      const double y_y = y_xi * xi_y + y_eta * eta_y + y_zeta * zeta_y;
      const double y_z = y_xi * xi_z + y_eta * eta_z + y_zeta * zeta_z;
      const double GX = xi_x * y_x + xi_y * y_y + xi_z * y_z;
      const double GE = eta_x * y_x + eta_y * y_y + eta_z * y_z;
      const double GZ = zeta_x * y_x + zeta_y * y_y + zeta_z * y_z;
      dfx[l] = GX;
      dfy[l] = GE;
      dfz[l] = GZ;
      CVx += diffusivity[l] * GX;
      CVy += diffusivity[l] * GE;
      CVz += diffusivity[l] * GZ;
      sGx += GX * tb->imw[l];
      sGy += GE * tb->imw[l];
      sGz += GZ * tb->imw[l];
      mw_tot += yk[l] * tb->imw[l];
      sum_rhoDkYk += diffusivity[l] * yk[l];
    }
    mw_tot = 1.0 / mw_tot;
    CVx -= mw_tot * sum_rhoDkYk * sGx;
    CVy -= mw_tot * sum_rhoDkYk * sGy;
    CVz -= mw_tot * sum_rhoDkYk * sGz;
    if (gradp) {
      const double p_xi = DXI(pv, 4), p_eta = DETA(pv, 4), p_zeta = z.zg[t + (4 + NS) * dsz];
      const double p_x = p_xi * xi_x + p_eta * eta_x + p_zeta * zeta_x;
      const double p_y = p_xi * xi_y + p_eta * eta_y + p_zeta * zeta_y;
      const double p_z = p_xi * xi_z + p_eta * eta_z + p_zeta * zeta_z;
      const double pp = pv[p + 4 * sz];
      const double gX = (xi_x * p_x + xi_y * p_y + xi_z * p_z) / pp;
      const double gE = (eta_x * p_x + eta_y * p_y + eta_z * p_z) / pp;
      const double gZ = (zeta_x * p_x + zeta_y * p_y + zeta_z * p_z) / pp;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const double coefficient = (mw_tot * tb->imw[l] - 1) * yk[l];
        dfx[l] += coefficient * gX;
        dfy[l] += coefficient * gE;
        dfz[l] += coefficient * gZ;
        CVx += coefficient * gX * diffusivity[l];
        CVy += coefficient * gE * diffusivity[l];
        CVz += coefficient * gZ * diffusivity[l];
      }
    }
    double h[NS];
    const double tm = pv[p + 5 * sz];
    thermo_h<M>(tb, tm, h);
    [[maybe_unused]] const double tve_m = kTT(M) ? z.tve[p] : tm;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      // compute_nonequilibrium_diffusion_enthalpy, src/Thermo.cuh:43-47, and the
      // E_ve the diffusion flux carries, src/ViscousScheme.cu:722-756
      double h_diff = h[l];
      [[maybe_unused]] double eve_l = 0;
      if constexpr (kTT(M)) {
        eve_l = ve_energy1<false>(tb, l, tve_m, nullptr);
        h_diff = h[l] - ve_energy1<false>(tb, l, tm, nullptr) + eve_l;
      }
      double dflux = diffusivity[l] * (dfx[l] - mw_tot * yk[l] * sGx) - yk[l] * CVx;
      fv[p + (5 + l) * sz] = dflux;
      f4 += h_diff * dflux;
      if constexpr (kTT(M)) eve_fx += eve_l * dflux;
      dflux = diffusivity[l] * (dfy[l] - mw_tot * yk[l] * sGy) - yk[l] * CVy;
      gv[p + (5 + l) * sz] = dflux;
      g4 += h_diff * dflux;
      if constexpr (kTT(M)) eve_fy += eve_l * dflux;
      dflux = diffusivity[l] * (dfz[l] - mw_tot * yk[l] * sGz) - yk[l] * CVz;
      hv[p + (5 + l) * sz] = dflux;
      h4 += h_diff * dflux;
      if constexpr (kTT(M)) eve_fz += eve_l * dflux;
    }
    if constexpr (kTT(M)) {
      fv[p + (5 + NS) * sz] = eve_fx;
      gv[p + (5 + NS) * sz] = eve_fy;
      hv[p + (5 + NS) * sz] = eve_fz;
    }
  }
  fv[p + 4 * sz] = f4;
  gv[p + 4 * sz] = g4;
  hv[p + 4 * sz] = h4;
}

// compute_viscous_flux_derivative<8>, src/ViscousScheme.cu:790-834:
// dq += J (dF/dxi + dG/deta + dH/dzeta).  A thread owns KC consecutive k of
// one (i,j) column and slides a nine-node window of each component's H along
// them, so H is read (KC + 8) / KC times instead of nine.
// STORE: the derivative is the stage's first contribution to dq and replaces
// it (a stage recorded without the separate clearing of dq); else it adds.
template <int NV, int KC, bool STORE>
__global__ __launch_bounds__(kBlock) void k_viscous_derivative(Zone z, long ncol) {
  const long t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= ncol) return;
  const int mx = z.mx, my = z.my, mz = z.mz;
  const int i = static_cast<int>(t % mx);
  const long r = t / mx;
  const int j = static_cast<int>(r % my);
  const int k0 = static_cast<int>(r / my) * KC;
  const int nk = min(KC, mz - k0);
  const long sz = z.sz, si = 1, sj = z.sx, sk = z.sxy, dsz = z.dsz, dk = static_cast<long>(mx) * my;
  const bool kl = z.bt_kl[i + j * mx] != 0, kr = z.bt_kr[i + j * mx] != 0;
  const long p0 = z.ix(i, j, k0), q0 = z.dx(i, j, k0);
  static_assert(KC <= 16, "four flag bits per node of a chunk");
  // the i / j boundary flags of the chunk's nodes, four bits per node, read once for all components
  unsigned long long fl = 0;
  __shared__ double jac_s[KC][kBlock];  // the chunk's J, each thread's own column
  for (int c = 0; c < nk; ++c) {
    const int k = k0 + c;
    const unsigned b = (z.bt_il[j + k * my] != 0 ? 1u : 0u) | (z.bt_ir[j + k * my] != 0 ? 2u : 0u) |
                       (z.bt_jl[i + k * mx] != 0 ? 4u : 0u) | (z.bt_jr[i + k * mx] != 0 ? 8u : 0u);
    fl |= static_cast<unsigned long long>(b) << (4 * c);
    jac_s[c][threadIdx.x] = z.jac[p0 + c * sk];
  }
  if constexpr (STORE) {
    // the continuity equation has no viscous term: its dq starts at 0 here
    for (int c = 0; c < nk; ++c) z.dq[q0 + c * dk] = 0.0;
  }
#pragma unroll 1
  for (int l = 1; l < NV; ++l) {
    // nodes k0-4 .. k0+nk+3 lie inside the padded extent (ngg >= 4)
    double w[9];
#pragma unroll
    for (int m = 0; m < 8; ++m) w[m] = z.hf[p0 + (m - 4) * sk + l * sz];
#pragma unroll 1
    for (int c = 0; c < nk; ++c) {
      const int k = k0 + c;
      const long p = p0 + c * sk;
      w[8] = z.hf[p + 4 * sk + l * sz];
      const unsigned b = static_cast<unsigned>(fl >> (4 * c));
      const bool il = (b & 1u) != 0, ir = (b & 2u) != 0, jl = (b & 4u) != 0, jr = (b & 8u) != 0;
      const double dxi = d8(z.ff, p + l * sz, si, i, mx, il, ir);
      const double deta = d8(z.gf, p + l * sz, sj, j, my, jl, jr);
      const double dzeta = d8g([&](int o) { return w[4 + o]; }, k, mz, kl, kr);
      if constexpr (STORE)
        z.dq[q0 + c * dk + l * dsz] = (dxi + deta + dzeta) * jac_s[c][threadIdx.x];
      else
        z.dq[q0 + c * dk + l * dsz] += (dxi + deta + dzeta) * jac_s[c][threadIdx.x];
#pragma unroll
      for (int m = 0; m < 8; ++m) w[m] = w[m + 1];
    }
  }
}

#undef DXI
#undef DETA
#undef DZETA

// compute_{f,g,h}v_2nd_order (one temperature), src/ViscousScheme.cu:836-1512:
// the viscous flux on the face between node P and P + e_DIR from first
// differences across the face and four-point central differences along the
// other axes, with face-averaged mu, kappa, rho*D, metrics and J grad(DIR).
// One thread per face (faces -1 .. n-1 along DIR); fFlux holds the faces.
template <int M, int DIR>
__global__ __launch_bounds__(kBlock) void k_visc2_flux(Zone z, const Tab *__restrict__ tb, double prandtl, int gradp,
                                                      long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  const int nx = z.mx + (DIR == 0), ny = z.my + (DIR == 1);
  int i = static_cast<int>(t % nx);
  const long r = t / nx;
  int j = static_cast<int>(r % ny);
  int k = static_cast<int>(r / ny);
  if constexpr (DIR == 0) i -= 1;
  if constexpr (DIR == 1) j -= 1;
  if constexpr (DIR == 2) k -= 1;
  const long sz = z.sz;
  const long st[3] = {1, z.sx, z.sxy};
  const long p = z.ix(i, j, k), pe = p + st[DIR];
  const double *pv = z.bv, *mt = z.metric;
  double m[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) m[a] = 0.5 * (mt[p + a * sz] + mt[pe + a * sz]);
  auto dd = [&](const double *arr, int l, int a) {
    const long o = static_cast<long>(l) * sz;
    if (a == DIR) return arr[pe + o] - arr[p + o];
    return 0.25 * (arr[p + st[a] + o] - arr[p - st[a] + o] + arr[pe + st[a] + o] - arr[pe - st[a] + o]);
  };
  auto grad = [&](const double *arr, int l, double &gx, double &gy, double &gz) {
    const double d0 = dd(arr, l, 0), d1 = dd(arr, l, 1), d2 = dd(arr, l, 2);
    gx = d0 * m[0] + d1 * m[3] + d2 * m[6];
    gy = d0 * m[1] + d1 * m[4] + d2 * m[7];
    gz = d0 * m[2] + d1 * m[5] + d2 * m[8];
  };
  double u_x, u_y, u_z, v_x, v_y, v_z, w_x, w_y, w_z, t_x, t_y, t_z;
  grad(pv, 1, u_x, u_y, u_z);
  grad(pv, 2, v_x, v_y, v_z);
  grad(pv, 3, w_x, w_y, w_z);
  grad(pv, 5, t_x, t_y, t_z);
  const double mul = 0.5 * (z.mul[p] + z.mul[pe]);
  const double tau_xx = mul * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  const double tau_yy = mul * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  const double tau_zz = mul * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const double tau_xy = mul * (u_y + v_x), tau_xz = mul * (u_z + w_x), tau_yz = mul * (v_z + w_y);
  const double kx = 0.5 * (mt[p + (3 * DIR) * sz] * z.jac[p] + mt[pe + (3 * DIR) * sz] * z.jac[pe]);
  const double ky = 0.5 * (mt[p + (3 * DIR + 1) * sz] * z.jac[p] + mt[pe + (3 * DIR + 1) * sz] * z.jac[pe]);
  const double kz = 0.5 * (mt[p + (3 * DIR + 2) * sz] * z.jac[p] + mt[pe + (3 * DIR + 2) * sz] * z.jac[pe]);
  double *fv = z.ff;
  const double f0 = kx * tau_xx + ky * tau_xy + kz * tau_xz;
  const double f1 = kx * tau_xy + ky * tau_yy + kz * tau_yz;
  const double f2 = kx * tau_xz + ky * tau_yz + kz * tau_zz;
  fv[p] = f0;
  fv[p + sz] = f1;
  fv[p + 2 * sz] = f2;
  const double um = 0.5 * (pv[p + sz] + pv[pe + sz]), vm = 0.5 * (pv[p + 2 * sz] + pv[pe + 2 * sz]),
               wm = 0.5 * (pv[p + 3 * sz] + pv[pe + 3 * sz]);
  double conductivity;
  if constexpr (NS > 0) {
    conductivity = 0.5 * (z.kappa[p] + z.kappa[pe]);
  } else {
    constexpr double cp_air = kGammaAir * kRair / (kGammaAir - 1);
    conductivity = mul / prandtl * cp_air;
  }
  double f3 = um * f0 + vm * f1 + wm * f2 + conductivity * (kx * t_x + ky * t_y + kz * t_z);
  if constexpr (NS > 0) {
    const double *y = z.sv;
    double diff[NS], yk[NS], force[NS];
    double sGw = 0, sDY = 0, CV = 0, mw_tot = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      yk[l] = 0.5 * (y[p + l * sz] + y[pe + l * sz]);
      diff[l] = 0.5 * (z.rhoD[p + l * sz] + z.rhoD[pe + l * sz]);
      double y_x, y_y, y_z;
      grad(y, l, y_x, y_y, y_z);
      const double G = kx * y_x + ky * y_y + kz * y_z;
      force[l] = G;
      CV += diff[l] * G;
      sGw += G * tb->imw[l];
      mw_tot += yk[l] * tb->imw[l];
      sDY += diff[l] * yk[l];
    }
    mw_tot = 1.0 / mw_tot;
    CV -= mw_tot * sDY * sGw;
    if (gradp) {
      double p_x, p_y, p_z;
      grad(pv, 4, p_x, p_y, p_z);
      const double gp = (kx * p_x + ky * p_y + kz * p_z) / (0.5 * (pv[pe + 4 * sz] + pv[p + 4 * sz]));
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        force[l] += (mw_tot * tb->imw[l] - 1) * yk[l] * gp;
        CV += (mw_tot * tb->imw[l] - 1) * yk[l] * gp * diff[l];
      }
    }
    double h[NS];
    const double tm = 0.5 * (pv[p + 5 * sz] + pv[pe + 5 * sz]);
    thermo_h<M>(tb, tm, h);
    [[maybe_unused]] double tve_m = tm, eve_m[NS];
    if constexpr (kTT(M)) {
      tve_m = 0.5 * (z.tve[p] + z.tve[pe]);
#pragma unroll
      for (int l = 0; l < NS; ++l) eve_m[l] = ve_energy1<false>(tb, l, tve_m, nullptr);
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double dflux = diff[l] * (force[l] - mw_tot * yk[l] * sGw) - yk[l] * CV;
      fv[p + (4 + l) * sz] = dflux;
      double h_diff = h[l];
      // compute_nonequilibrium_diffusion_enthalpy, src/Thermo.cuh:43-47
      if constexpr (kTT(M)) h_diff = h[l] - ve_energy1<false>(tb, l, tm, nullptr) + eve_m[l];
      f3 += h_diff * dflux;
    }
    if constexpr (kTT(M)) {
      // kappa_ve grad Tve conduction and the E_ve flux, src/ViscousScheme.cu:1040-1055
      double tve_x, tve_y, tve_z;
      grad(z.tve, 0, tve_x, tve_y, tve_z);
      const double k_ve = 0.5 * (z.kappa_ve[p] + z.kappa_ve[pe]);
      const double cond = k_ve * (kx * tve_x + ky * tve_y + kz * tve_z);
      f3 += cond;
      double eve_flux = cond;
#pragma unroll
      for (int l = 0; l < NS; ++l) eve_flux += eve_m[l] * fv[p + (4 + l) * sz];
      fv[p + (4 + NS) * sz] = eve_flux;
    }
  }
  fv[p + 3 * sz] = f3;
}

// compute_d{F,G,H}v_d{x,y,z}, src/ViscousScheme.cu:86-138: dq[1..] += F(face) - F(face - 1).
template <int NV, int DIR>
__global__ __launch_bounds__(kBlock) void k_visc2_diff(Zone z, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz;
  const long back = DIR == 0 ? 1 : (DIR == 1 ? z.sx : z.sxy);
#pragma unroll
  for (int l = 1; l < NV; ++l) z.dq[t + l * z.dsz] += z.ff[p + (l - 1) * sz] - z.ff[p - back + (l - 1) * sz];
}

// One thread per (face point, layer) of a boundary face; layer 0 is the face
// node itself, layers 1..g its ghosts.  On i-faces the layer index runs
// fastest so that neighbouring lanes touch neighbouring addresses; on j/k
// faces i runs fastest.
struct FaceSite {
  int i, j, k, layer;
};
DEV bool face_site(const Zone &z, const corefl_boundary &b, long t, FaceSite &f) {
  const int n0 = b.range_end[0] - b.range_start[0] + 1, n1 = b.range_end[1] - b.range_start[1] + 1;
  const int n2 = b.range_end[2] - b.range_start[2] + 1;
  const int nl = z.ngg + 1;
  if (t >= static_cast<long>(n0) * n1 * n2 * nl) return false;
  int a0, a1, a2;
  if (b.face == 0) {
    f.layer = static_cast<int>(t % nl);
    const long r = t / nl;
    a0 = 0;
    a1 = static_cast<int>(r % n1);
    a2 = static_cast<int>(r / n1);
  } else {
    a0 = static_cast<int>(t % n0);
    const long r = t / n0;
    const int nm = b.face == 1 ? n2 : n1;  // the other tangential extent
    const int o = static_cast<int>(r % nm);
    f.layer = static_cast<int>(r / nm);
    a1 = b.face == 1 ? 0 : o;
    a2 = b.face == 1 ? o : 0;
  }
  f.i = b.range_start[0] + a0;
  f.j = b.range_start[1] + a1;
  f.k = b.range_start[2] + a2;
  return true;
}

// periodic_exchange(task 1), src/DataCommunication.cu:665-703: average the
// coincident face nodes' viscous fluxes, copy ghosts from beyond the opposite face.
// The divergence reads ghosts across a face only for the flux along the face's
// normal (fFlux beyond an x face, gFlux beyond y, hFlux beyond z), so only
// that array's ghost layers are copied; the face nodes average all three.
__global__ void k_periodic_flux(Zone z, corefl_boundary b, int nv) {
  FaceSite f;
  if (!face_site(z, b, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, f)) return;
  int o[3] = {f.i, f.j, f.k};
  const int nn[3] = {z.mx, z.my, z.mz};
  o[b.face] = b.direction < 0 ? nn[b.face] - 1 : 0;
  int d[3] = {0, 0, 0};
  d[b.face] = b.direction;
  const long sz = z.sz;
  double *arrs[3] = {z.ff, z.gf, z.hf};
  // all loads of a plane set before its stores (the planes never overlap; the
  // compiler would otherwise serialise each load behind the previous store)
  double va[kMaxNv], vb[kMaxNv];
  if (f.layer == 0) {
    const long p = z.ix(f.i, f.j, f.k), po = z.ix(o[0], o[1], o[2]);
    for (int a = 0; a < 3; ++a) {
      double *fa = arrs[a];
#pragma unroll
      for (int l = 0; l < kMaxNv; ++l)
        if (l < nv) {
          va[l] = fa[p + l * sz];
          vb[l] = fa[po + l * sz];
        }
#pragma unroll
      for (int l = 0; l < kMaxNv; ++l)
        if (l < nv) {
          const double ave = 0.5 * (va[l] + vb[l]);
          fa[p + l * sz] = ave;
          fa[po + l * sz] = ave;
        }
    }
  } else {
    const int g = f.layer;
    const long pg = z.ix(f.i + g * d[0], f.j + g * d[1], f.k + g * d[2]);
    const long pi = z.ix(o[0] + g * d[0], o[1] + g * d[1], o[2] + g * d[2]);
    double *fa = arrs[b.face];
#pragma unroll
    for (int l = 0; l < kMaxNv; ++l)
      if (l < nv) va[l] = fa[pi + l * sz];
#pragma unroll
    for (int l = 0; l < kMaxNv; ++l)
      if (l < nv) fa[pg + l * sz] = va[l];
  }
}

// The two periodic faces of one axis in one launch (the faces' own launches
// in sequence average the same coincident nodes twice, the second time to the
// same values, and copy disjoint ghost layers): threads over the low face's
// sites; layer 0 averages the coincident pair once, layer g fills both ghosts.
__global__ void k_periodic_flux_pair(Zone z, corefl_boundary lo, int nv) {
  FaceSite f;
  if (!face_site(z, lo, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, f)) return;
  const int a = lo.face;
  const int n = a == 0 ? z.mx : (a == 1 ? z.my : z.mz);
  int o[3] = {f.i, f.j, f.k};
  o[a] = n - 1;
  const long sz = z.sz;
  double *arrs[3] = {z.ff, z.gf, z.hf};
  double va[kMaxNv], vb[kMaxNv];
  if (f.layer == 0) {
    const long p = z.ix(f.i, f.j, f.k), po = z.ix(o[0], o[1], o[2]);
    for (int c = 0; c < 3; ++c) {
      double *fa = arrs[c];
#pragma unroll
      for (int l = 0; l < kMaxNv; ++l)
        if (l < nv) {
          va[l] = fa[p + l * sz];
          vb[l] = fa[po + l * sz];
        }
#pragma unroll
      for (int l = 0; l < kMaxNv; ++l)
        if (l < nv) {
          const double ave = 0.5 * (va[l] + vb[l]);
          fa[p + l * sz] = ave;
          fa[po + l * sz] = ave;
        }
    }
  } else {
    const int g = f.layer;
    int lg[3] = {f.i, f.j, f.k}, li[3] = {f.i, f.j, f.k}, hg[3] = {f.i, f.j, f.k}, hi[3] = {f.i, f.j, f.k};
    lg[a] = -g;
    li[a] = n - 1 - g;
    hg[a] = n - 1 + g;
    hi[a] = g;
    double *fa = arrs[a];
    const long plg = z.ix(lg[0], lg[1], lg[2]), pli = z.ix(li[0], li[1], li[2]);
    const long phg = z.ix(hg[0], hg[1], hg[2]), phi = z.ix(hi[0], hi[1], hi[2]);
#pragma unroll
    for (int l = 0; l < kMaxNv; ++l)
      if (l < nv) {
        va[l] = fa[pli + l * sz];
        vb[l] = fa[phi + l * sz];
      }
#pragma unroll
    for (int l = 0; l < kMaxNv; ++l)
      if (l < nv) {
        fa[plg + l * sz] = va[l];
        fa[phg + l * sz] = vb[l];
      }
  }
}

// apply_outflow (zero-gradient branch), src/BoundCond.cu:2204-2349: the
// ghost layers outside face node p (step d along the outward normal) copy
// its state.
// (ghost layer g of a face node: the layers are independent of each other)
template <int M>
DEV void bc_outflow_layer(const Zone &z, const Tab *__restrict__ tb, int i, int j, int k, const int *d, int g) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long p = z.ix(i, j, k), sz = z.sz;
  const long pg = z.ix(i + g * d[0], j + g * d[1], k + g * d[2]);
  double b6[6], sc[kNsc(M) > 0 ? kNsc(M) : 1];
#pragma unroll
  for (int l = 0; l < 6; ++l) b6[l] = z.bv[p + l * sz];
#pragma unroll
  for (int l = 0; l < kNsc(M); ++l) sc[l] = z.sv[p + l * sz];
  [[maybe_unused]] const double tve = kTT(M) ? z.tve[p] : 0.0;
#pragma unroll
  for (int l = 0; l < 6; ++l) z.bv[pg + l * sz] = b6[l];
#pragma unroll
  for (int l = 0; l < kNsc(M); ++l) z.sv[pg + l * sz] = sc[l];
  if constexpr (kTT(M)) z.tve[pg] = tve > 0.0 ? tve : b6[5];
  cv_from_bv_point<M>(z, tb, pg);
}

template <int M>
DEV void bc_outflow_point(const Zone &z, const Tab *__restrict__ tb, int i, int j, int k, const int *d) {
  for (int g = 1; g <= z.ngg; ++g) bc_outflow_layer<M>(z, tb, i, j, k, d, g);
}

template <int M>
__global__ void k_bc_outflow(Zone z, const Tab *__restrict__ tb, corefl_boundary b) {
  const int n0 = b.range_end[0] - b.range_start[0] + 1, n1 = b.range_end[1] - b.range_start[1] + 1;
  const int n2 = b.range_end[2] - b.range_start[2] + 1;
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<long>(n0) * n1 * n2) return;
  const int i = b.range_start[0] + static_cast<int>(t % n0);
  const int j = b.range_start[1] + static_cast<int>((t / n0) % n1);
  const int k = b.range_start[2] + static_cast<int>(t / (static_cast<long>(n0) * n1));
  int d[3] = {0, 0, 0};
  d[b.face] = b.direction;
  bc_outflow_point<M>(z, tb, i, j, k, d);
}

// Face point t of a boundary face (i fastest).
DEV bool face_point(const corefl_boundary &b, long t, int &i, int &j, int &k) {
  const int n0 = b.range_end[0] - b.range_start[0] + 1, n1 = b.range_end[1] - b.range_start[1] + 1;
  const int n2 = b.range_end[2] - b.range_start[2] + 1;
  if (t >= static_cast<long>(n0) * n1 * n2) return false;
  i = b.range_start[0] + static_cast<int>(t % n0);
  j = b.range_start[1] + static_cast<int>((t / n0) % n1);
  k = b.range_start[2] + static_cast<int>(t / (static_cast<long>(n0) * n1));
  return true;
}

template <int M>
DEV double gas_constant_at(const Zone &z, const Tab *__restrict__ tb, long p) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  if constexpr (NS == 0) {
    return kRair;
  } else {
    double R = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) R += z.sv[p + l * z.sz] * tb->gas_const[l];
    return R;
  }
}

// apply_symmetry, src/BoundCond.cu:2130-2202: the face node keeps the
// tangential velocity of its interior neighbour and its rho, p, T and
// scalars; ghost layer g mirrors interior layer g with the normal velocity
// reversed.
template <int M>
DEV void bc_symmetry_layer(const Zone &z, const Tab *__restrict__ tb, int i, int j, int k, int face, int direction,
                           int layer) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  int d[3] = {0, 0, 0};
  d[face] = direction;
  const long sz = z.sz, p = z.ix(i, j, k);
  double kx = z.metric[p + (3 * face) * sz], ky = z.metric[p + (3 * face + 1) * sz],
         kz = z.metric[p + (3 * face + 2) * sz];
  const double km = sqrt(kx * kx + ky * ky + kz * kz);
  kx /= km;
  ky /= km;
  kz /= km;
  if (layer == 0) {
    const long pin = z.ix(i - d[0], j - d[1], k - d[2]);
    const double u1 = z.bv[pin + sz], v1 = z.bv[pin + 2 * sz], w1 = z.bv[pin + 3 * sz];
    const double u_k = kx * u1 + ky * v1 + kz * w1;
    z.bv[p + sz] = u1 - kx * u_k;
    z.bv[p + 2 * sz] = v1 - ky * u_k;
    z.bv[p + 3 * sz] = w1 - kz * u_k;
    z.bv[p] = z.bv[pin];
    z.bv[p + 4 * sz] = z.bv[pin + 4 * sz];
    z.bv[p + 5 * sz] = z.bv[pin + 5 * sz];
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) z.sv[p + l * sz] = z.sv[pin + l * sz];
    if constexpr (kTT(M)) z.tve[p] = z.tve[pin];
    cv_from_bv_point<M>(z, tb, p);
    return;
  }
  const int g = layer;
  const long pg = z.ix(i + g * d[0], j + g * d[1], k + g * d[2]);
  const long pi = z.ix(i - g * d[0], j - g * d[1], k - g * d[2]);
  z.bv[pg] = z.bv[pi];
  const double u = z.bv[pi + sz], v = z.bv[pi + 2 * sz], w = z.bv[pi + 3 * sz];
  const double u_k = kx * u + ky * v + kz * w;
  z.bv[pg + sz] = u - 2 * u_k * kx;
  z.bv[pg + 2 * sz] = v - 2 * u_k * ky;
  z.bv[pg + 3 * sz] = w - 2 * u_k * kz;
  z.bv[pg + 4 * sz] = z.bv[pi + 4 * sz];
  z.bv[pg + 5 * sz] = z.bv[pi + 5 * sz];
#pragma unroll
  for (int l = 0; l < kNsc(M); ++l) z.sv[pg + l * sz] = z.sv[pi + l * sz];
  if constexpr (kTT(M)) z.tve[pg] = z.tve[pi];
  cv_from_bv_point<M>(z, tb, pg);
}

// The face node (layer 0) and the ghost layers read only interior nodes, so
// they are independent of each other.
template <int M>
DEV void bc_symmetry_point(const Zone &z, const Tab *__restrict__ tb, int i, int j, int k, int face, int direction) {
  for (int layer = 0; layer <= z.ngg; ++layer) bc_symmetry_layer<M>(z, tb, i, j, k, face, direction, layer);
}

template <int M>
__global__ void k_bc_symmetry(Zone z, const Tab *__restrict__ tb, corefl_boundary b) {
  int i, j, k;
  if (!face_point(b, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, i, j, k)) return;
  bc_symmetry_point<M>(z, tb, i, j, k, b.face, b.direction);
}

// apply_inflow, constant inflow without fluctuations, src/BoundCond.cu:2810-2997:
// the face node takes the inflow state, or with inflow_sub (unless the flow
// at the face enters supersonically, Mn < -1) the characteristic
// combination of the inflow state with the interior neighbour; ghosts take
// the face state, or with inflow_sub its reflection about the face.
template <int M>
__global__ void k_bc_inflow(Zone z, const Tab *__restrict__ tb, corefl_boundary b, corefl_bc_condition c) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  int i, j, k;
  if (!face_point(b, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, i, j, k)) return;
  int d[3] = {0, 0, 0};
  d[b.face] = b.direction;
  const long sz = z.sz, p = z.ix(i, j, k);
  double density = c.density, u = c.u, v = c.v, w = c.w, pr = c.pressure, T = c.temperature;
  if (c.inflow_sub) {
    double nx = z.metric[p + (3 * b.face) * sz], ny = z.metric[p + (3 * b.face + 1) * sz],
           nz = z.metric[p + (3 * b.face + 2) * sz];
    const double gn = b.direction / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= gn;
    ny *= gn;
    nz *= gn;
    const double c0 = NS > 0 ? z.acoustic[p] : sqrt(kGammaAir * kRair * z.bv[p + 5 * sz]);
    const double Mn = (z.bv[p + sz] * nx + z.bv[p + 2 * sz] * ny + z.bv[p + 3 * sz] * nz) / c0;
    if (!(Mn < -1.0)) {
      const long p1 = z.ix(i - d[0], j - d[1], k - d[2]);
      const double c1 = NS > 0 ? z.acoustic[p1] : sqrt(kGammaAir * kRair * z.bv[p1 + 5 * sz]);
      const double rho1 = z.bv[p1], p1v = z.bv[p1 + 4 * sz];
      pr = 0.5 * (p1v + c.pressure -
                  rho1 * c1 * ((c.u - z.bv[p1 + sz]) * nx + (c.v - z.bv[p1 + 2 * sz]) * ny +
                               (c.w - z.bv[p1 + 3 * sz]) * nz));
      density = c.density + (pr - c.pressure) / (c1 * c1);
      u = c.u - (pr - c.pressure) / (rho1 * c1) * nx;
      v = c.v - (pr - c.pressure) / (rho1 * c1) * ny;
      w = c.w - (pr - c.pressure) / (rho1 * c1) * nz;
      T = NS > 0 ? pr * c.mw / (density * kRu) : pr / (density * kRair);
    }
  }
  z.bv[p] = density;
  z.bv[p + sz] = u;
  z.bv[p + 2 * sz] = v;
  z.bv[p + 3 * sz] = w;
  z.bv[p + 4 * sz] = pr;
  z.bv[p + 5 * sz] = T;
#pragma unroll
  for (int l = 0; l < kNsc(M); ++l) z.sv[p + l * sz] = c.sv[l];
  cv_from_bv_point<M>(z, tb, p);
  for (int g = 1; g <= z.ngg; ++g) {
    const long pg = z.ix(i + g * d[0], j + g * d[1], k + g * d[2]);
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) z.sv[pg + l * sz] = c.sv[l];
    if (c.inflow_sub) {
      const long pi = z.ix(i - g * d[0], j - g * d[1], k - g * d[2]);
      double rho2 = 2 * density - z.bv[pi], p2 = 2 * pr - z.bv[pi + 4 * sz];
      if (p2 <= 0.1 * pr) p2 = 0.1 * pr;
      if (rho2 <= 0.1 * density) rho2 = 0.1 * density;
      z.bv[pg] = rho2;
      z.bv[pg + sz] = 2 * u - z.bv[pi + sz];
      z.bv[pg + 2 * sz] = 2 * v - z.bv[pi + 2 * sz];
      z.bv[pg + 3 * sz] = 2 * w - z.bv[pi + 3 * sz];
      z.bv[pg + 4 * sz] = p2;
      z.bv[pg + 5 * sz] = NS > 0 ? p2 * c.mw / (rho2 * kRu) : p2 / (rho2 * kRair);
    } else {
      z.bv[pg] = density;
      z.bv[pg + sz] = u;
      z.bv[pg + 2 * sz] = v;
      z.bv[pg + 3 * sz] = w;
      z.bv[pg + 4 * sz] = pr;
      z.bv[pg + 5 * sz] = T;
    }
    cv_from_bv_point<M>(z, tb, pg);
  }
}

// project_to_interior in apply_wall, src/BoundCond.cu:3070-3087: a reference
// node lying on another physical face moves one node inward along that axis.
DEV void project_to_interior(const Zone &z, int *ix) {
  if (ix[0] <= 0 && z.bt_il[ix[1] + ix[2] * z.my] != 0)
    ix[0] = 1;
  else if (ix[0] >= z.mx - 1 && z.bt_ir[ix[1] + ix[2] * z.my] != 0)
    ix[0] = z.mx - 2;
  if (ix[1] <= 0 && z.bt_jl[ix[0] + ix[2] * z.mx] != 0)
    ix[1] = 1;
  else if (ix[1] >= z.my - 1 && z.bt_jr[ix[0] + ix[2] * z.mx] != 0)
    ix[1] = z.my - 2;
  if (z.mz > 1) {
    if (ix[2] <= 0 && z.bt_kl[ix[0] + ix[1] * z.mx] != 0)
      ix[2] = 1;
    else if (ix[2] >= z.mz - 1 && z.bt_kr[ix[0] + ix[1] * z.mx] != 0)
      ix[2] = z.mz - 2;
  }
}

// apply_wall for a resting, non-catalytic wall, isothermal or adiabatic,
// src/BoundCond.cu:3055-3387: the face node is no-slip with the wall
// temperature and the pressure of its interior neighbour; ghost layer g
// reflects the velocity of interior layer g and takes T_g = 2 T_w - T_i
// (isothermal; T_w where that falls to 0.1 T_w or below) or T_i.
template <int M>
__global__ void k_bc_wall(Zone z, const Tab *__restrict__ tb, corefl_boundary b, corefl_bc_condition c) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  int i, j, k;
  if (!face_point(b, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, i, j, k)) return;
  int d[3] = {0, 0, 0};
  d[b.face] = b.direction;
  const long sz = z.sz, p = z.ix(i, j, k);
  int ix[3] = {i - d[0], j - d[1], k - d[2]};
  project_to_interior(z, ix);
  const long pin = z.ix(ix[0], ix[1], ix[2]);
  const double t_wall = c.thermal_type == 0 ? c.wall_temperature : z.bv[pin + 5 * sz];
  [[maybe_unused]] double tve_wall = t_wall;
  if constexpr (kTT(M)) {
    // Tve of the wall: the interior's (adiabatic) or the condition's, else T_w (src/BoundCond.cu:3167-3177)
    const double tve_in = z.tve[pin] > 0.0 ? z.tve[pin] : z.bv[pin + 5 * sz];
    tve_wall = c.thermal_type == 1 ? tve_in : (c.temperature_ve > 0.0 ? c.temperature_ve : t_wall);
  }
  const double pr = z.bv[pin + 4 * sz];
#pragma unroll
  for (int l = 0; l < NS; ++l) z.sv[p + l * sz] = z.sv[pin + l * sz];
  z.bv[p] = pr / (t_wall * gas_constant_at<M>(z, tb, p));
  z.bv[p + sz] = 0;
  z.bv[p + 2 * sz] = 0;
  z.bv[p + 3 * sz] = 0;
  z.bv[p + 4 * sz] = pr;
  z.bv[p + 5 * sz] = t_wall;
  if constexpr (kTT(M)) {
    double y[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) y[l] = z.sv[p + l * sz];
    z.tve[p] = tve_wall;
    z.sv[p + NS * sz] = mix_ve_energy<M>(tb, tve_wall, y, nullptr);
  }
  cv_from_bv_point<M>(z, tb, p);
  for (int g = 1; g <= z.ngg; ++g) {
    int in[3] = {i - g * d[0], j - g * d[1], k - g * d[2]};
    project_to_interior(z, in);
    const long pi = z.ix(in[0], in[1], in[2]);
    const long pg = z.ix(i + g * d[0], j + g * d[1], k + g * d[2]);
    const double p_i = z.bv[pi + 4 * sz], t_i = z.bv[pi + 5 * sz];
    double t_g = t_i;
    if (c.thermal_type == 0) {
      t_g = 2 * t_wall - t_i;
      // the floor is 0.1 T_w in the two-temperature build and T_w otherwise
      if (t_g <= 0.1 * t_wall) t_g = kTT(M) ? 0.1 * t_wall : t_wall;
    }
#pragma unroll
    for (int l = 0; l < NS; ++l) z.sv[pg + l * sz] = z.sv[pi + l * sz];
    if constexpr (kTT(M)) {
      const double tve_i = z.tve[pi] > 0.0 ? z.tve[pi] : t_i;
      double tve_g = tve_i;
      if (c.thermal_type == 0) {
        tve_g = 2 * tve_wall - tve_i;
        const double floor_ve = 0.1 * fmax(tve_wall, 1.0);
        if (tve_g <= floor_ve) tve_g = floor_ve;
      }
      double y[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) y[l] = z.sv[pg + l * sz];
      z.tve[pg] = tve_g;
      z.sv[pg + NS * sz] = mix_ve_energy<M>(tb, tve_g, y, nullptr);
    }
    z.bv[pg] = p_i / (t_g * gas_constant_at<M>(z, tb, pi));
    z.bv[pg + sz] = -z.bv[pi + sz];
    z.bv[pg + 2 * sz] = -z.bv[pi + 2 * sz];
    z.bv[pg + 3 * sz] = -z.bv[pi + 3 * sz];
    z.bv[pg + 4 * sz] = p_i;
    z.bv[pg + 5 * sz] = t_g;
    cv_from_bv_point<M>(z, tb, pg);
  }
}

// limit_flow, src/TimeAdvanceFunc.cuh:129-232.  A point with a NaN or an
// out-of-bounds rho, u, v, w or p takes the mean of the good points of its
// 3x3x3 neighbourhood (clipped to the bounds when there is none).  Every
// repair is computed from the state before any repair (fFlux[0..4] and the
// flag gFlux[0] hold them) and applied by a second kernel.
DEV bool flow_bad(const Zone &z, long p, const Run &r) {
  for (int l = 0; l < 5; ++l) {
    const double v = z.bv[p + l * z.sz];
    if (isnan(v) || v < r.ll[l] || v > r.ul[l]) return true;
  }
  return false;
}

__global__ __launch_bounds__(kBlock) void k_limit_flow_scan(Zone z, Run r, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz;
  const bool bad = flow_bad(z, p, r);
  z.gf[p] = bad ? 1.0 : 0.0;
  if (!bad) return;
  double acc[5] = {0, 0, 0, 0, 0};
  int kn = 0;
  for (int ka = -1; ka < 2; ++ka) {
    const int k1 = k + ka;
    if (k1 < 0 || k1 >= z.mz) continue;
    for (int ja = -1; ja < 2; ++ja) {
      const int j1 = j + ja;
      if (j1 < 0 || j1 >= z.my) continue;
      for (int ia = -1; ia < 2; ++ia) {
        const int i1 = i + ia;
        if (i1 < 0 || i1 >= z.mx) continue;
        const long p1 = z.ix(i1, j1, k1);
        if (flow_bad(z, p1, r)) continue;
#pragma unroll
        for (int l = 0; l < 5; ++l) acc[l] += z.bv[p1 + l * sz];
        ++kn;
      }
    }
  }
  if (kn > 0) {
    const double inv = 1.0 / kn;
#pragma unroll
    for (int l = 0; l < 5; ++l) acc[l] *= inv;
  } else {
#pragma unroll
    for (int l = 0; l < 5; ++l) acc[l] = fmin(fmax(z.bv[p + l * sz], r.ll[l]), r.ul[l]);
  }
#pragma unroll
  for (int l = 0; l < 5; ++l) z.ff[p + l * sz] = acc[l];
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_limit_flow_apply(Zone z, const Tab *__restrict__ tb, long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz;
  if (z.gf[p] == 0.0) return;
#pragma unroll
  for (int l = 0; l < 5; ++l) z.bv[p + l * sz] = z.ff[p + l * sz];
  z.bv[p + 5 * sz] = z.ff[p + 4 * sz] / (z.ff[p] * gas_constant_at<M>(z, tb, p));
  cv_from_bv_point<M>(z, tb, p);
}

// Loop position t of an inner face -> node in this block (idx) and in the
// target block (it), src/DataCommunication.cu:43-52.  n[0] runs fastest.
DEV void inner_nodes(const InnerDev &f, long t, int *idx, int *it) {
  int n[3];
  n[0] = static_cast<int>(t % f.n_point[0]);
  const long r = t / f.n_point[0];
  n[1] = static_cast<int>(r % f.n_point[1]);
  n[2] = static_cast<int>(r / f.n_point[1]);
  int d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    d[a] = f.loop_dir[a] * n[a];
    idx[a] = f.range_start[a] + d[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) it[a] = f.target_start[a] + f.target_loop_dir[a] * d[f.src_tar[a]];
}

// inner_communication, src/DataCommunication.cu:37-124: the coincident node
// takes the average of both blocks (the neighbour's values where this side
// moves and the neighbour is at rest; nothing where this side is at rest);
// ghosts take the neighbour's interior; cv follows bv on every node written.
// Coincident-node threads write the target only on its face nodes, ghost
// threads read it only inside, so the threads of one launch are independent.
template <int M>
__global__ void k_inner_comm(Zone z, Zone tz, const Tab *__restrict__ tb, InnerDev f, int n_scalar, long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int idx[3], it[3];
  inner_nodes(f, t, idx, it);
  const long p = z.ix(idx[0], idx[1], idx[2]), q = tz.ix(it[0], it[1], it[2]);
  const long sz = z.sz, tsz = tz.sz;
  const int face_dir = f.direction > 0 ? f.range_start[f.face] : f.range_end[f.face];
  if (idx[f.face] == face_dir) {
    const double u_this = norm3(z.bv[p + sz], z.bv[p + 2 * sz], z.bv[p + 3 * sz]);
    if (fabs(u_this) > 1e-15) {
      const double u_that = norm3(tz.bv[q + tsz], tz.bv[q + 2 * tsz], tz.bv[q + 3 * tsz]);
      if (fabs(u_that) < 1e-15) {
        for (int l = 0; l < 6; ++l) z.bv[p + l * sz] = tz.bv[q + l * tsz];
        for (int l = 0; l < n_scalar; ++l) z.sv[p + l * sz] = tz.sv[q + l * tsz];
        if constexpr (kTT(M)) z.tve[p] = tz.tve[q];
      } else {
        for (int l = 0; l < 6; ++l) {
          const double ave = 0.5 * (tz.bv[q + l * tsz] + z.bv[p + l * sz]);
          z.bv[p + l * sz] = ave;
          tz.bv[q + l * tsz] = ave;
        }
        for (int l = 0; l < n_scalar; ++l) {
          const double ave = 0.5 * (tz.sv[q + l * tsz] + z.sv[p + l * sz]);
          z.sv[p + l * sz] = ave;
          tz.sv[q + l * tsz] = ave;
        }
        if constexpr (kTT(M)) {
          const double ave = 0.5 * (tz.tve[q] + z.tve[p]);
          z.tve[p] = ave;
          tz.tve[q] = ave;
        }
        cv_from_bv_point<M>(tz, tb, q);
      }
    }
  } else {
    for (int l = 0; l < 6; ++l) z.bv[p + l * sz] = tz.bv[q + l * tsz];
    for (int l = 0; l < n_scalar; ++l) z.sv[p + l * sz] = tz.sv[q + l * tsz];
    if constexpr (kTT(M)) z.tve[p] = tz.tve[q];
  }
  cv_from_bv_point<M>(z, tb, p);
}

// inner_exchange(task 1), src/DataCommunication.cu:489-527: coincident nodes
// average f/g/hFlux between the blocks, ghost nodes take the neighbour's.
__global__ void k_inner_flux(Zone z, Zone tz, InnerDev f, int nv, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  int idx[3], it[3];
  inner_nodes(f, t, idx, it);
  const long p = z.ix(idx[0], idx[1], idx[2]), q = tz.ix(it[0], it[1], it[2]);
  const long sz = z.sz, tsz = tz.sz;
  const int face_dir = f.direction > 0 ? f.range_start[f.face] : f.range_end[f.face];
  double *a[3] = {z.ff, z.gf, z.hf};
  double *b[3] = {tz.ff, tz.gf, tz.hf};
  if (idx[f.face] == face_dir) {
    for (int l = 0; l < nv; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double ave = 0.5 * (a[c][p + l * sz] + b[c][q + l * tsz]);
        a[c][p + l * sz] = ave;
        b[c][q + l * tsz] = ave;
      }
  } else {
    for (int l = 0; l < nv; ++l)
#pragma unroll
      for (int c = 0; c < 3; ++c) a[c][p + l * sz] = b[c][q + l * tsz];
  }
}

// apply_periodic, src/BoundCond.cu:3389-3440: average the coincident face
// nodes (bv, sv) and recompute cv there; ghosts copy the nodes beyond the
// opposite face.  Ghost sources are interior nodes no thread of the launch writes.
template <int M>
__global__ void k_bc_periodic(Zone z, const Tab *__restrict__ tb, corefl_boundary b) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  FaceSite f;
  if (!face_site(z, b, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, f)) return;
  int o[3] = {f.i, f.j, f.k};
  const int nn[3] = {z.mx, z.my, z.mz};
  o[b.face] = b.direction < 0 ? nn[b.face] - 1 : 0;
  int d[3] = {0, 0, 0};
  d[b.face] = b.direction;
  const long sz = z.sz;
  // all loads before the stores (see k_periodic_flux)
  constexpr int NSC = kNsc(M) > 0 ? kNsc(M) : 1;
  double a6[6], asc[NSC];
  if (f.layer == 0) {
    const long p = z.ix(f.i, f.j, f.k), po = z.ix(o[0], o[1], o[2]);
    double b6[6], bsc[NSC];
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      a6[l] = z.bv[p + l * sz];
      b6[l] = z.bv[po + l * sz];
    }
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) {
      asc[l] = z.sv[p + l * sz];
      bsc[l] = z.sv[po + l * sz];
    }
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      const double ave = 0.5 * (a6[l] + b6[l]);
      z.bv[p + l * sz] = ave;
      z.bv[po + l * sz] = ave;
    }
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) {
      const double ave = 0.5 * (asc[l] + bsc[l]);
      z.sv[p + l * sz] = ave;
      z.sv[po + l * sz] = ave;
    }
    cv_from_bv_point<M>(z, tb, p);
  } else {
    const int g = f.layer;
    const long pg = z.ix(f.i + g * d[0], f.j + g * d[1], f.k + g * d[2]);
    const long pi = z.ix(o[0] + g * d[0], o[1] + g * d[1], o[2] + g * d[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) a6[l] = z.bv[pi + l * sz];
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) asc[l] = z.sv[pi + l * sz];
#pragma unroll
    for (int l = 0; l < 6; ++l) z.bv[pg + l * sz] = a6[l];
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) z.sv[pg + l * sz] = asc[l];
    cv_from_bv_point<M>(z, tb, pg);
  }
}

// The two periodic faces of one axis in one launch (see k_periodic_flux_pair):
// layer 0 averages the coincident nodes once and forms both nodes'
// conservative variables, layer g fills the low and the high ghost.
template <int M>
__global__ void k_bc_periodic_pair(Zone z, const Tab *__restrict__ tb, corefl_boundary lo) {
  FaceSite f;
  if (!face_site(z, lo, blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x, f)) return;
  const int a = lo.face;
  const int n = a == 0 ? z.mx : (a == 1 ? z.my : z.mz);
  const long sz = z.sz;
  constexpr int NSC = kNsc(M) > 0 ? kNsc(M) : 1;
  double a6[6], asc[NSC], b6[6], bsc[NSC];
  if (f.layer == 0) {
    int o[3] = {f.i, f.j, f.k};
    o[a] = n - 1;
    const long p = z.ix(f.i, f.j, f.k), po = z.ix(o[0], o[1], o[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      a6[l] = z.bv[p + l * sz];
      b6[l] = z.bv[po + l * sz];
    }
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) {
      asc[l] = z.sv[p + l * sz];
      bsc[l] = z.sv[po + l * sz];
    }
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      const double ave = 0.5 * (a6[l] + b6[l]);
      z.bv[p + l * sz] = ave;
      z.bv[po + l * sz] = ave;
    }
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) {
      const double ave = 0.5 * (asc[l] + bsc[l]);
      z.sv[p + l * sz] = ave;
      z.sv[po + l * sz] = ave;
    }
    cv_from_bv_point<M>(z, tb, p);
    cv_from_bv_point<M>(z, tb, po);
  } else {
    const int g = f.layer;
    int lg[3] = {f.i, f.j, f.k}, li[3] = {f.i, f.j, f.k}, hg[3] = {f.i, f.j, f.k}, hi[3] = {f.i, f.j, f.k};
    lg[a] = -g;
    li[a] = n - 1 - g;
    hg[a] = n - 1 + g;
    hi[a] = g;
    const long plg = z.ix(lg[0], lg[1], lg[2]), pli = z.ix(li[0], li[1], li[2]);
    const long phg = z.ix(hg[0], hg[1], hg[2]), phi = z.ix(hi[0], hi[1], hi[2]);
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      a6[l] = z.bv[pli + l * sz];
      b6[l] = z.bv[phi + l * sz];
    }
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) {
      asc[l] = z.sv[pli + l * sz];
      bsc[l] = z.sv[phi + l * sz];
    }
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      z.bv[plg + l * sz] = a6[l];
      z.bv[phg + l * sz] = b6[l];
    }
#pragma unroll
    for (int l = 0; l < kNsc(M); ++l) {
      z.sv[plg + l * sz] = asc[l];
      z.sv[phg + l * sz] = bsc[l];
    }
    cv_from_bv_point<M>(z, tb, plg);
    cv_from_bv_point<M>(z, tb, phg);
  }
}

// eliminate_k_gradient, src/FieldOperation.cu:96-121
__global__ void k_k_gradient(Zone z, int n_scalar, int nv) {
  const long nxp = z.mx + 2 * z.ngg, nyp = z.my + 2 * z.ngg;
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= nxp * nyp) return;
  const int i = static_cast<int>(t % nxp) - z.ngg, j = static_cast<int>(t / nxp) - z.ngg;
  const long p0 = z.ix(i, j, 0), sz = z.sz;
  for (int k = 1; k <= z.ngg; ++k) {
    const long pp = z.ix(i, j, k), pm = z.ix(i, j, -k);
    for (int l = 0; l < 6; ++l) z.bv[pp + l * sz] = z.bv[pm + l * sz] = z.bv[p0 + l * sz];
    for (int l = 0; l < n_scalar; ++l) z.sv[pp + l * sz] = z.sv[pm + l * sz] = z.sv[p0 + l * sz];
    for (int l = 0; l < nv; ++l) z.cv[pp + l * sz] = z.cv[pm + l * sz] = z.cv[p0 + l * sz];
  }
}

// local_time_step (1T), src/TimeAdvanceFunc.cuh:28-127, with a per-block min
template <int M>
__global__ __launch_bounds__(kBlock) void k_local_dt(Zone z, const Tab *__restrict__ tb, double cfl, double prandtl,
                                                     int viscous, int reaction, double *block_min, long n) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  __shared__ double red[kBlock / 64];
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  double dt = 1e+6;
  if (t < n) {
    int i, j, k;
    interior_ijk(z, t, i, j, k);
    const long p = z.ix(i, j, k), sz = z.sz;
    const double *mt = z.metric;
    const int dim = z.mz == 1 ? 2 : 3;
    const double grad_xi = norm3(mt[p], mt[p + sz], mt[p + 2 * sz]);
    const double grad_eta = norm3(mt[p + 3 * sz], mt[p + 4 * sz], mt[p + 5 * sz]);
    const double grad_zeta = norm3(mt[p + 6 * sz], mt[p + 7 * sz], mt[p + 8 * sz]);
    const double u = z.bv[p + sz], v = z.bv[p + 2 * sz], w = z.bv[p + 3 * sz];
    const double U = u * mt[p] + v * mt[p + sz] + w * mt[p + 2 * sz];
    const double V = u * mt[p + 3 * sz] + v * mt[p + 4 * sz] + w * mt[p + 5 * sz];
    const double W = u * mt[p + 6 * sz] + v * mt[p + 7 * sz] + w * mt[p + 8 * sz];
    double c;
    if constexpr (NS > 0)
      c = z.acoustic[p];
    else
      c = sqrt(kGammaAir * kRair * z.bv[p + 5 * sz]);
    double msr = fmax(fabs(U) + c * grad_xi, fabs(V) + c * grad_eta);
    if (dim == 3) msr = fmax(msr, fabs(W) + c * grad_zeta);
    if (viscous > 0) {
      double max_length = fmax(grad_xi, grad_eta);
      if (dim == 3) max_length = fmax(max_length, grad_zeta);
      const double iRho = 1.0 / z.bv[p];
      double mdv;
      if constexpr (NS > 0) {
        double cpl[NS];
        thermo_cp<M>(tb, z.bv[p + 5 * sz], cpl);
        double cp = 0, max_rhoD = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          cp += cpl[l] * z.sv[p + l * sz];
          max_rhoD = fmax(max_rhoD, z.rhoD[p + l * sz]);
        }
        mdv = z.mul[p] * iRho;
        mdv = fmax(mdv, z.kappa[p] * iRho * z.gam[p] / cp);
        mdv = fmax(mdv, max_rhoD * iRho);
        if constexpr (kTT(M)) {
          // src/TimeAdvanceFunc.cu:75-85
          double y[NS], cv_ve = 0.0;
#pragma unroll
          for (int l = 0; l < NS; ++l) y[l] = z.sv[p + l * sz];
          const double tve = z.tve[p] > 0.0 ? z.tve[p] : fmax(z.bv[p + 5 * sz], 1.0);
          mix_ve_energy<M>(tb, tve, y, &cv_ve);
          mdv = fmax(mdv, z.kappa_ve[p] * iRho / fmax(cv_ve, 1e-8));
        }
      } else {
        mdv = fmax(kGammaAir, 4.0 / 3.0) * iRho;
        mdv *= z.mul[p] / prandtl;
      }
      msr = fmax(msr, max_length * max_length * mdv);
    }
    dt = cfl / msr;
    if constexpr (NS > 0) {
      if (reaction) {
        const double reaction_dt = cfl * z.rts[p];
        if (reaction_dt > 1e-20 && reaction_dt < 1e-5) dt = fmin(dt, reaction_dt);
      }
    }
    z.dt_local[t] = dt;
  }
  // wave64 reduction then across the block's waves
  for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_down(dt, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = dt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w2 = 1; w2 < kBlock / 64; ++w2) m = fmin(m, red[w2]);
    block_min[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(kBlock) void k_min_reduce(double *data, long n, double *out) {
  __shared__ double red[kBlock / 64];
  double m = 1e+6;
  for (long t = threadIdx.x; t < n; t += blockDim.x) m = fmin(m, data[t]);
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_down(m, off, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w2 = 1; w2 < kBlock / 64; ++w2) r = fmin(r, red[w2]);
    *out = r;
  }
}

// The step's dt as the minimum of the blocks' minima (global_time_step,
// src/TimeAdvanceFunc.cu:114-139); one thread, exact in any order.
__global__ void k_min_combine(const double *v, int n, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double m = 1e+6;
  for (int b = 0; b < n; ++b) m = fmin(m, v[b]);
  *out = m;
}

inline unsigned blocks_for(long n, int b = kBlock) { return static_cast<unsigned>((n + b - 1) / b); }
inline long padded_points(const Zone &z) {
  return static_cast<long>(z.mx + 2 * z.ngg) * (z.my + 2 * z.ngg) * (z.mz + 2 * z.ngg);
}
inline long interior_points(const Zone &z) { return static_cast<long>(z.mx) * z.my * z.mz; }
inline long face_points(const corefl_boundary &b) {
  return static_cast<long>(b.range_end[0] - b.range_start[0] + 1) * (b.range_end[1] - b.range_start[1] + 1) *
         (b.range_end[2] - b.range_start[2] + 1);
}

// Dispatch over the compiled thermochemical models (kNs/kNasa9/kTT codes):
// 0 Air, 5 and 9 species with NASA-7 thermodynamics, 5 species with NASA-9
// (37) and 5 species with NASA-9 and two temperatures (101).
#define COREFL_DISPATCH(model, CALL) \
  switch (model) {                  \
    case 0: {                       \
      constexpr int MV = 0;         \
      CALL;                         \
    } break;                        \
    case 5: {                       \
      constexpr int MV = 5;         \
      CALL;                         \
    } break;                        \
    case 9: {                       \
      constexpr int MV = 9;         \
      CALL;                         \
    } break;                        \
    case 37: {                      \
      constexpr int MV = 37;        \
      CALL;                         \
    } break;                        \
    case 101: {                     \
      constexpr int MV = 101;       \
      CALL;                         \
    } break;                        \
    default:                        \
      return hipErrorInvalidValue;  \
  }

// The same over the species mixtures only.
#define COREFL_DISPATCH_MIX(model, CALL) \
  switch (model) {                      \
    case 5: {                           \
      constexpr int MV = 5;             \
      CALL;                             \
    } break;                            \
    case 9: {                           \
      constexpr int MV = 9;             \
      CALL;                             \
    } break;                            \
    case 37: {                          \
      constexpr int MV = 37;            \
      CALL;                             \
    } break;                            \
    case 101: {                         \
      constexpr int MV = 101;           \
      CALL;                             \
    } break;                            \
    default:                            \
      return hipErrorInvalidValue;      \
  }

// Tile shapes of the characteristic sweeps (TX points across, TS along the
// sweep; a block computes TS - 1 faces).  Every shape computes the same faces
// with the same arithmetic, so the choice changes the time, not the result;
// the driver times the shapes on each block and keeps the fastest.
template <int M, bool W7, int DIR, int TX, int TS>
void launch_weno_shape(const Zone &z, const Tab *tb, double eps, const WenoOpt &o, hipStream_t s) {
  dim3 grid;
  if constexpr (DIR == 0)
    grid = dim3((z.mx + TS - 2) / (TS - 1), (z.my + TX - 1) / TX, z.mz);
  else if constexpr (DIR == 1)
    grid = dim3((z.mx + TX - 1) / TX, (z.my + TS - 2) / (TS - 1), z.mz);
  else
    grid = dim3((z.mx + TX - 1) / TX, z.my, (z.mz + TS - 2) / (TS - 1));
  hipLaunchKernelGGL((k_weno<M, W7, DIR, TX, TS>), grid, dim3(TX * TS), 0, s, z, tb, eps, o);
}

template <int M, bool W7>
hipError_t launch_weno_t(const Zone &z, const Tab *tb, int dir, double eps, const WenoOpt &o, hipStream_t s,
                         int shape) {
  if (dir == 0) {
    // shape 1: two lines of 31 new faces, for rows whose node count leaves the 63-face tile mostly idle
    // in its last block (256 nodes: 5 tiles of 63 for 256 faces against 9 of 31 per line pair)
    if (shape == 1)
      launch_weno_shape<M, W7, 0, 2, 32>(z, tb, eps, o, s);
    else
      launch_weno_shape<M, W7, 0, 1, 64>(z, tb, eps, o, s);
  } else if (dir == 1) {
    if (shape == 1)
      launch_weno_shape<M, W7, 1, 8, 32>(z, tb, eps, o, s);
    else
      launch_weno_shape<M, W7, 1, 8, 16>(z, tb, eps, o, s);
  } else {
    if (shape == 1)
      launch_weno_shape<M, W7, 2, 8, 32>(z, tb, eps, o, s);
    else
      launch_weno_shape<M, W7, 2, 8, 16>(z, tb, eps, o, s);
  }
  return hipGetLastError();
}

template <int M, bool W7>
hipError_t launch_weno_cp_t(const Zone &z, int dir, double eps, const WenoOpt &o, hipStream_t s) {
  if (dir == 0) {
    constexpr int TX = 1, TS = 64;
    dim3 grid((z.mx + TS - 2) / (TS - 1), (z.my + TX - 1) / TX, z.mz);
    hipLaunchKernelGGL((k_weno_cp<M, W7, 0, TX, TS>), grid, dim3(TX * TS), 0, s, z, eps, o);
  } else if (dir == 1) {
    constexpr int TX = 8, TS = 16;
    dim3 grid((z.mx + TX - 1) / TX, (z.my + TS - 2) / (TS - 1), z.mz);
    hipLaunchKernelGGL((k_weno_cp<M, W7, 1, TX, TS>), grid, dim3(TX * TS), 0, s, z, eps, o);
  } else {
    constexpr int TX = 8, TS = 16;
    dim3 grid((z.mx + TX - 1) / TX, z.my, (z.mz + TS - 2) / (TS - 1));
    hipLaunchKernelGGL((k_weno_cp<M, W7, 2, TX, TS>), grid, dim3(TX * TS), 0, s, z, eps, o);
  }
  return hipGetLastError();
}

template <int M, int RW>
hipError_t launch_pv_t(const Zone &z, const Tab *tb, int dir, const PvArgs &a, hipStream_t s) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  if (dir == 0) {
    constexpr int TX = 1, TS = 64;
    dim3 grid((z.mx + TS - 2) / (TS - 1), (z.my + TX - 1) / TX, z.mz);
    hipLaunchKernelGGL((k_pv<M, 0, TX, TS, RW>), grid, dim3(TX * TS), 0, s, z, tb, a);
  } else if (dir == 1) {
    constexpr int TX = 8, TS = 16;
    dim3 grid((z.mx + TX - 1) / TX, (z.my + TS - 2) / (TS - 1), z.mz);
    hipLaunchKernelGGL((k_pv<M, 1, TX, TS, RW>), grid, dim3(TX * TS), 0, s, z, tb, a);
  } else {
    constexpr int TX = 8, TS = 16;
    dim3 grid((z.mx + TX - 1) / TX, z.my, (z.mz + TS - 2) / (TS - 1));
    hipLaunchKernelGGL((k_pv<M, 2, TX, TS, RW>), grid, dim3(TX * TS), 0, s, z, tb, a);
  }
  return hipGetLastError();
}

}  // namespace

bool model_supported(int model) { return model == 0 || model == 5 || model == 9 || model == 37 || model == 101; }

hipError_t launch_inviscid_pv(const Zone &z, const Run &r, const Tab *tb, int dir, hipStream_t s) {
  if (r.scheme < 1 || r.scheme > 4 || (r.scheme == 1 && r.mixture)) return hipErrorInvalidValue;
  const double vr2 = r.v_ref * r.v_ref;
  const PvArgs a{r.scheme, r.recon, r.efix, z.mz == 1 ? 2 : 3, 1e-40 * r.rho_ref * r.rho_ref, 1e-40 * vr2,
                 1e-40 * (r.rho_ref * vr2) * (r.rho_ref * vr2)};
  if (r.recon == 5) {
    COREFL_DISPATCH(r.model, return (launch_pv_t<MV, 4>(z, tb, dir, a, s)));
  } else if (r.recon == 4) {
    COREFL_DISPATCH(r.model, return (launch_pv_t<MV, 3>(z, tb, dir, a, s)));
  } else {
    COREFL_DISPATCH(r.model, return (launch_pv_t<MV, 2>(z, tb, dir, a, s)));
  }
  return hipSuccess;
}

hipError_t launch_fill(double *p, long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_scaled_copy(double *dst, const double *src, double a, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scaled_copy, dim3(blocks_for(n)), dim3(kBlock), 0, s, dst, src, a, n);
  return hipGetLastError();
}

hipError_t launch_copy(double *dst, const double *src, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(n)), dim3(kBlock), 0, s, dst, src, n);
  return hipGetLastError();
}

hipError_t launch_cv_from_bv(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  const long n = padded_points(z);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_cv_from_bv<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, n));
  return hipGetLastError();
}

hipError_t launch_properties(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  const long n = padded_points(z);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_properties<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb,
                                        r.viscous_order == 2 ? 1 : 0, n));
  return hipGetLastError();
}

hipError_t launch_update_rk(const Zone &z, const Run &r, const Tab *tb, double dt, int rk, hipStream_t s,
                            const double *dtp, const double *dtc, bool save_qn) {
  const long n = interior_points(z);
  if (save_qn) {
    COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_update_rk<MV, false, true>), dim3(blocks_for(n)), dim3(kBlock), 0,
                                                s, z, tb, dt, dtp, dtc, rk, n));
  } else {
    COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_update_rk<MV>), dim3(blocks_for(n)), dim3(kBlock), 0,
                                                s, z, tb, dt, dtp, dtc, rk, n));
  }
  return hipGetLastError();
}

hipError_t launch_chemistry(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  if (!r.mixture) return hipErrorInvalidValue;
  const long n = interior_points(z);
  COREFL_DISPATCH_MIX(r.model, hipLaunchKernelGGL((k_chemistry<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, n));
  return hipGetLastError();
}

hipError_t launch_vt_relaxation(const Zone &z, const Run &r, const Tab *tb, double dt_stage, hipStream_t s,
                                const double *dtp, double scale) {
  if (!r.tt || (!dtp && dt_stage <= 0.0)) return hipSuccess;
  const long n = interior_points(z);
  switch (r.model) {
    case 101:
      hipLaunchKernelGGL((k_vt_relax<101>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, dt_stage, dtp, scale, n);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_inviscid(const Zone &z, const Run &r, const Tab *tb, int dir, hipStream_t s, int shape) {
  const bool w7 = r.scheme == 72 || r.scheme == 71, charw = r.scheme == 52 || r.scheme == 72;
  if (!charw && r.scheme != 51 && r.scheme != 71) return hipErrorInvalidValue;
  if (r.hybrid && !z.sensor) return hipErrorInvalidValue;
  WenoOpt o{};
  o.sensor = r.hybrid ? z.sensor : nullptr;
  o.threshold = r.sensor_threshold;
  o.pp = r.pp;
  o.pp_dt = r.pp_dt;
  o.alpha = z.mz == 1 ? 0.5 : 1.0 / 3.0;
  o.rho_ref = r.rho_ref;
  o.v_ref = r.v_ref;
  if (charw) {
    if (w7) {
      COREFL_DISPATCH(r.model, return (launch_weno_t<MV, true>(z, tb, dir, r.eps_ref, o, s, shape)));
    } else {
      COREFL_DISPATCH(r.model, return (launch_weno_t<MV, false>(z, tb, dir, r.eps_ref, o, s, shape)));
    }
  } else {
    if (w7) {
      COREFL_DISPATCH(r.model, return (launch_weno_cp_t<MV, true>(z, dir, r.eps_ref, o, s)));
    } else {
      COREFL_DISPATCH(r.model, return (launch_weno_cp_t<MV, false>(z, dir, r.eps_ref, o, s)));
    }
  }
  return hipSuccess;
}

hipError_t launch_reactors(const ReactorPool &p, int model, const Tab *tb, int max_steps, hipStream_t s) {
  const unsigned nb = blocks_for(p.cap, kReactorBlock);
  COREFL_DISPATCH_MIX(model, hipLaunchKernelGGL((k_reactors<MV>), dim3(nb), dim3(kReactorBlock), 0, s, p, tb,
                                                max_steps));
  return hipGetLastError();
}

hipError_t launch_reactors_wave(const ReactorPool &p, int model, const Tab *tb, int max_steps, int n_slots,
                                hipStream_t s) {
  if (n_slots <= 0) return hipSuccess;
  COREFL_DISPATCH_MIX(model, hipLaunchKernelGGL((k_reactors_wave<MV>), dim3(n_slots), dim3(kWave), 0, s, p, tb,
                                                max_steps, n_slots));
  return hipGetLastError();
}

hipError_t launch_reactor_admit(const ReactorPool &p, int model, const Tab *tb, const double *req, int n_req,
                                hipStream_t s) {
  if (n_req <= 0) return hipSuccess;
  const unsigned nb = blocks_for(n_req, 64);
  COREFL_DISPATCH_MIX(model, hipLaunchKernelGGL((k_reactor_admit<MV>), dim3(nb), dim3(64), 0, s, p, tb, req, n_req));
  return hipGetLastError();
}

hipError_t launch_shock_sensor(const Zone &z, const Run &r, hipStream_t s) {
  if (!z.sensor) return hipErrorInvalidValue;
  const long n = interior_points(z);
  hipLaunchKernelGGL(k_shock_sensor, dim3(blocks_for(n)), dim3(kBlock), 0, s, z, z.sensor, r.sensor_type,
                     r.sensor_eps, n);
  return hipGetLastError();
}

hipError_t launch_viscous_flux(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  constexpr int KC = 16;
  const long ncol = static_cast<long>(z.mx) * z.my * ((z.mz + KC - 1) / KC);
  const int ns = r.mixture ? r.ns : 0;
  // zeta-derivative fields: u, v, w, T, Y_l, (p with the pressure-gradient term), (Tve)
  const int nf = 4 + ns + (r.gradp && ns > 0 ? 1 : 0);
  const int tve_f = r.tt ? nf : -1;
  hipLaunchKernelGGL((k_zeta_grad<KC>), dim3(blocks_for(ncol)), dim3(kBlock), 0, s, z, ns, nf + (r.tt ? 1 : 0),
                     tve_f, ncol);
  const long n = interior_points(z);
  COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_viscous_flux<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                              tb, r.prandtl, r.gradp, tve_f, n));
  return hipGetLastError();
}

template <int M, int DIR>
hipError_t launch_visc2_dir(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  const long nf = static_cast<long>(z.mx + (DIR == 0)) * (z.my + (DIR == 1)) * (z.mz + (DIR == 2));
  hipLaunchKernelGGL((k_visc2_flux<M, DIR>), dim3(blocks_for(nf)), dim3(kBlock), 0, s, z, tb, r.prandtl, r.gradp,
                     nf);
  const long n = interior_points(z);
  hipLaunchKernelGGL((k_visc2_diff<kNv(M), DIR>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, n);
  return hipGetLastError();
}

template <int M>
hipError_t launch_visc2_t(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  [[maybe_unused]] constexpr int NS = kNs(M);
  hipError_t e = launch_visc2_dir<M, 0>(z, r, tb, s);
  if (e == hipSuccess) e = launch_visc2_dir<M, 1>(z, r, tb, s);
  if (e == hipSuccess && z.mz > 1) e = launch_visc2_dir<M, 2>(z, r, tb, s);
  return e;
}

hipError_t launch_viscous2(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  COREFL_DISPATCH(r.model, return launch_visc2_t<MV>(z, r, tb, s));
  return hipSuccess;
}

hipError_t launch_viscous_derivative(const Zone &z, const Run &r, hipStream_t s, bool store) {
  constexpr int KC = 16;
  const long ncol = static_cast<long>(z.mx) * z.my * ((z.mz + KC - 1) / KC);
  if (store) {
    COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_viscous_derivative<kNv(MV), KC, true>),
                                                dim3(blocks_for(ncol)), dim3(kBlock), 0, s, z, ncol));
  } else {
    COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_viscous_derivative<kNv(MV), KC, false>),
                                                dim3(blocks_for(ncol)), dim3(kBlock), 0, s, z, ncol));
  }
  return hipGetLastError();
}

hipError_t launch_periodic_flux(const Zone &z, const Run &r, const corefl_boundary &b, hipStream_t s) {
  const long n = face_points(b) * (z.ngg + 1);
  hipLaunchKernelGGL(k_periodic_flux, dim3(blocks_for(n)), dim3(kBlock), 0, s, z, b, r.nv);
  return hipGetLastError();
}

hipError_t launch_bc_outflow(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b, hipStream_t s) {
  const long n = face_points(b);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_bc_outflow<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, b));
  return hipGetLastError();
}

hipError_t launch_bc_symmetry(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b, hipStream_t s) {
  const long n = face_points(b);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_bc_symmetry<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, b));
  return hipGetLastError();
}

hipError_t launch_bc_inflow(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b,
                            const corefl_bc_condition &c, hipStream_t s) {
  const long n = face_points(b);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_bc_inflow<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, b, c));
  return hipGetLastError();
}

hipError_t launch_bc_wall(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b,
                          const corefl_bc_condition &c, hipStream_t s) {
  const long n = face_points(b);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_bc_wall<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, b, c));
  return hipGetLastError();
}

hipError_t launch_bc_periodic(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b, hipStream_t s) {
  const long n = face_points(b) * (z.ngg + 1);
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_bc_periodic<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tb, b));
  return hipGetLastError();
}

hipError_t launch_bc_periodic_pair(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &lo,
                                  hipStream_t s) {
  const long n = face_points(lo) * (z.ngg + 1);
  COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_bc_periodic_pair<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                              tb, lo));
  return hipGetLastError();
}

hipError_t launch_periodic_flux_pair(const Zone &z, const Run &r, const corefl_boundary &lo, hipStream_t s) {
  const long n = face_points(lo) * (z.ngg + 1);
  hipLaunchKernelGGL(k_periodic_flux_pair, dim3(blocks_for(n)), dim3(kBlock), 0, s, z, lo, r.nv);
  return hipGetLastError();
}

hipError_t launch_inner_comm(const Zone &z, const Zone &tz, const Run &r, const Tab *tb, const InnerDev &f,
                             hipStream_t s) {
  const long n = static_cast<long>(f.n_point[0]) * f.n_point[1] * f.n_point[2];
  COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_inner_comm<MV>), dim3(blocks_for(n)), dim3(kBlock),
                                                             0, s, z, tz, tb, f, r.n_scalar, n));
  return hipGetLastError();
}

hipError_t launch_inner_flux(const Zone &z, const Zone &tz, const Run &r, const InnerDev &f, hipStream_t s) {
  const long n = static_cast<long>(f.n_point[0]) * f.n_point[1] * f.n_point[2];
  hipLaunchKernelGGL(k_inner_flux, dim3(blocks_for(n)), dim3(kBlock), 0, s, z, tz, f, r.nv, n);
  return hipGetLastError();
}

hipError_t launch_limit_flow(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  const long n = interior_points(z);
  hipLaunchKernelGGL(k_limit_flow_scan, dim3(blocks_for(n)), dim3(kBlock), 0, s, z, r, n);
  COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_limit_flow_apply<MV>), dim3(blocks_for(n)),
                                                             dim3(kBlock), 0, s, z, tb, n));
  return hipGetLastError();
}

hipError_t launch_k_gradient(const Zone &z, const Run &r, hipStream_t s) {
  const long n = static_cast<long>(z.mx + 2 * z.ngg) * (z.my + 2 * z.ngg);
  hipLaunchKernelGGL(k_k_gradient, dim3(blocks_for(n)), dim3(kBlock), 0, s, z, r.n_scalar, r.nv);
  return hipGetLastError();
}

hipError_t launch_min_combine(const double *v, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_min_combine, dim3(1), dim3(64), 0, s, v, n, out);
  return hipGetLastError();
}

hipError_t launch_local_dt(const Zone &z, const Run &r, const Tab *tb, double *dmin, hipStream_t s) {
  const long n = interior_points(z);
  const unsigned nb = blocks_for(n);
  // per-block minima go to the tail of dt_local's allocation (sized n + nb by the driver)
  double *scratch = z.dt_local + n;
  COREFL_DISPATCH(r.model,
                     hipLaunchKernelGGL((k_local_dt<MV>), dim3(nb), dim3(kBlock), 0, s, z, tb, r.cfl, r.prandtl,
                                        r.viscous_order, r.reaction, scratch, n));
  hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(kBlock), 0, s, scratch, static_cast<long>(nb), dmin);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Multi-block launches of a recorded multi-block step: one launch covers the
// same operator on every block (blockIdx.y = block or face item), with the
// zones in device memory.  Each block's points get the arithmetic of the
// one-block kernels; blocks do not read each other here.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_zero_dq_multi(const Zone *__restrict__ zs, int nv) {
  const Zone &z = zs[blockIdx.y];
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t < z.dsz * nv) z.dq[t] = 0.0;
}

template <int M>
__global__ __launch_bounds__(kBlock) void k_bc_outflow_multi(const Zone *__restrict__ zs,
                                                             const BcItem *__restrict__ items,
                                                             const Tab *__restrict__ tb) {
  const BcItem &it = items[blockIdx.y];
  const corefl_boundary &b = it.b;
  const int n0 = b.range_end[0] - b.range_start[0] + 1, n1 = b.range_end[1] - b.range_start[1] + 1;
  const int n2 = b.range_end[2] - b.range_start[2] + 1;
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<long>(n0) * n1 * n2) return;
  const int i = b.range_start[0] + static_cast<int>(t % n0);
  const int j = b.range_start[1] + static_cast<int>((t / n0) % n1);
  const int k = b.range_start[2] + static_cast<int>(t / (static_cast<long>(n0) * n1));
  int d[3] = {0, 0, 0};
  d[b.face] = b.direction;
  bc_outflow_point<M>(zs[it.blk], tb, i, j, k, d);
}

hipError_t launch_zero_dq_multi(const Zone *zs, int nb, long max_n, int nv, hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_dq_multi, dim3(blocks_for(max_n * nv), nb), dim3(kBlock), 0, s, zs, nv);
  return hipGetLastError();
}

hipError_t launch_update_rk_multi(const Zone *zs, int nb, long max_n, const Run &r, const Tab *tb, const double *dtp,
                                  int rk, hipStream_t s, bool save_qn) {
  if (nb <= 0) return hipSuccess;
  if (save_qn) {
    COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_update_rk_multi<MV, true>), dim3(blocks_for(max_n), nb),
                                                dim3(kBlock), 0, s, zs, tb, dtp, rk));
  } else {
    COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_update_rk_multi<MV, false>), dim3(blocks_for(max_n), nb),
                                                dim3(kBlock), 0, s, zs, tb, dtp, rk));
  }
  return hipGetLastError();
}

hipError_t launch_properties_multi(const Zone *zs, int nb, long max_padded, const Run &r, const Tab *tb,
                                   hipStream_t s) {
  if (nb <= 0) return hipSuccess;
  COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_properties_multi<MV>), dim3(blocks_for(max_padded), nb),
                                              dim3(kBlock), 0, s, zs, tb, r.viscous_order == 2 ? 1 : 0));
  return hipGetLastError();
}

hipError_t launch_bc_outflow_multi(const Zone *zs, const BcItem *items, int n_items, long max_face, const Run &r,
                                   const Tab *tb, hipStream_t s) {
  if (n_items <= 0) return hipSuccess;
  COREFL_DISPATCH(r.model, hipLaunchKernelGGL((k_bc_outflow_multi<MV>), dim3(blocks_for(max_face), n_items),
                                              dim3(kBlock), 0, s, zs, items, tb));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched 1-D shock tubes (request server; no counterpart in the reference,
// which runs one grid per process).  The tubes lie end to end in one padded
// row, each with its own ghost layers on both sides, and the row is a Zone
// with my = mz = 1 whose y and z strides are 0.  The x sweep, chemistry,
// update and property kernels above run over the whole row in one launch
// each; a face between two tubes only feeds ghost points, whose updates the
// boundary conditions overwrite, so a tube's result does not depend on its
// neighbours.  Each tube takes its own CFL step: the RK update reads the step
// of the tube owning each point through a point -> slot map (points of no
// active tube step by 0).
// ---------------------------------------------------------------------------
#define COREFL_DISPATCH_1T(model, CALL) \
  switch (model) {                      \
    case 5: {                           \
      constexpr int MV = 5;             \
      CALL;                             \
    } break;                            \
    case 9: {                           \
      constexpr int MV = 9;             \
      CALL;                             \
    } break;                            \
    case 37: {                          \
      constexpr int MV = 37;            \
      CALL;                             \
    } break;                            \
    default:                            \
      return hipErrorInvalidValue;      \
  }

__global__ void k_tube_clear(int *map, const int *ranges, long map_n) {
  const int first = ranges[2 * blockIdx.x], count = ranges[2 * blockIdx.x + 1];
  for (int e = threadIdx.x; e < count; e += blockDim.x) {
    const long t = static_cast<long>(first) + e;
    if (t >= 0 && t < map_n) map[t] = -1;
  }
}

// One block per admission record: initial state (left state on nodes
// [0, n_left), right state on the rest; ghosts take their side's state until
// the boundary conditions fill them), metrics of the uniform lattice (1 / dx
// on the diagonal, jac = dx^3), the point -> slot map and the slot.  The
// record carries every derived number (density, 1 / dx, jac), computed once
// on the host.
template <int M>
__global__ void k_tube_admit(Zone z, TubePool P, const double *__restrict__ rec, int *map, long map_n) {
  constexpr int NS = kNs(M);
  const double *r = rec + static_cast<long>(blockIdx.x) * kTubeRecord;
  const int slot = static_cast<int>(r[0]), start = static_cast<int>(r[1]), n = static_cast<int>(r[2]);
  const int n_left = static_cast<int>(r[3]);
  const int g = z.ngg;
  const long sz = z.sz;
  for (int e = threadIdx.x; e < n + 2 * g; e += blockDim.x) {
    const int t = start - g + e, node = e - g;
    const int side = min(max(node, 0), n - 1) < n_left ? 0 : 1;
    const long p = z.ix(t, 0, 0);
    z.bv[p] = r[14 + side];
    z.bv[p + sz] = r[12 + side];
    z.bv[p + 2 * sz] = 0.0;
    z.bv[p + 3 * sz] = 0.0;
    z.bv[p + 4 * sz] = r[10 + side];
    z.bv[p + 5 * sz] = r[8 + side];
    const double *y = r + 19 + side * COREFL_MAX_SPEC;
    for (int l = 0; l < NS; ++l) z.sv[p + l * sz] = y[l];
    for (int m = 0; m < 9; ++m) z.metric[p + m * sz] = (m == 0 || m == 4 || m == 8) ? r[16] : 0.0;
    z.jac[p] = r[17];
    // the chemical time scale enters the first CFL step: a fresh tube starts without one, also in
    // row points an earlier tube used
    z.rts[p] = 0.0;
    if (t >= 0 && t < map_n) map[t] = (node >= 0 && node < n) ? slot : -1;
  }
  if (threadIdx.x == 0) {
    P.start[slot] = start;
    P.n[slot] = n;
    P.bc_l[slot] = static_cast<int>(r[6]);
    P.bc_r[slot] = static_cast<int>(r[7]);
    P.time[slot] = 0.0;
    P.t_end[slot] = r[5];
    P.dt[slot] = 0.0;
    P.last[slot] = 0;
    P.steps[slot] = 0;
    P.max_steps[slot] = static_cast<long>(r[18]);
    P.status[slot] = 1;
  }
}

// Conservative variables of the admitted tubes' nodes (after their boundary
// conditions, as the one-block initialisation does).
template <int M>
__global__ void k_tube_cv(Zone z, const Tab *__restrict__ tb, const double *__restrict__ rec) {
  const double *r = rec + static_cast<long>(blockIdx.x) * kTubeRecord;
  const int start = static_cast<int>(r[1]), n = static_cast<int>(r[2]);
  for (int e = threadIdx.x; e < n; e += blockDim.x) cv_from_bv_point<M>(z, tb, z.ix(start + e, 0, 0));
}

// Both ends of every active tube: outflow (zero gradient) or a reflecting
// end (symmetry), with the one-block kernels' layer functions; one thread per
// tube end and layer (the face node is layer 0).
template <int M>
__global__ void k_tube_bc(Zone z, TubePool P, const Tab *__restrict__ tb, int n_slots) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  const int nl = z.ngg + 1;
  const int slot = static_cast<int>(t / (2 * nl)), side = static_cast<int>((t / nl) & 1);
  const int layer = static_cast<int>(t % nl);
  if (slot >= n_slots || P.status[slot] != 1) return;
  const int dir = side ? 1 : -1;
  const int i = side ? P.start[slot] + P.n[slot] - 1 : P.start[slot];
  const int kind = side ? P.bc_r[slot] : P.bc_l[slot];
  if (kind == COREFL_BC_SYMMETRY) {
    bc_symmetry_layer<M>(z, tb, i, 0, 0, 0, dir, layer);
  } else if (layer > 0) {
    const int d[3] = {dir, 0, 0};
    bc_outflow_layer<M>(z, tb, i, 0, 0, d, layer);
  }
}

// One block per slot: the tube's CFL step (the minimum of dt_local over its
// nodes, global_time_step, src/TimeAdvanceFunc.cu:114-139), cut to reach
// t_end exactly on the last step.  A node whose local step is not finite and
// positive (fmin would drop a NaN) fails the tube: its step becomes 0, so its
// state stays as it is, and the host reports it.
// fixed_dt > 0: every tube steps by it (the last step cut to its t_end) instead of its CFL step.
__global__ __launch_bounds__(kBlock) void k_tube_dt(Zone z, TubePool P, double fixed_dt) {
  __shared__ double red[kBlock / 64];
  __shared__ int bad_w[kBlock / 64];
  const int slot = blockIdx.x;
  if (P.status[slot] != 1) return;
  const int start = P.start[slot], n = P.n[slot];
  double m = 1e+6;
  int bad = 0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const double d = z.dt_local[start + e];
    bad |= (d > 0 && d < INFINITY) ? 0 : 1;
    m = fmin(m, d);
  }
  for (int off = 32; off > 0; off >>= 1) {
    m = fmin(m, __shfl_down(m, off, 64));
    bad |= __shfl_down(bad, off, 64);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[wv] = m;
    bad_w[wv] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w2 = 1; w2 < kBlock / 64; ++w2) {
      m = fmin(m, red[w2]);
      bad |= bad_w[w2];
    }
    m = fmin(m, red[0]);
    bad |= bad_w[0];
    if (fixed_dt > 0) m = fixed_dt;
    if (bad || !(P.time[slot] < INFINITY)) {
      P.dt[slot] = 0.0;
      P.last[slot] = 0;
      P.status[slot] = kSlotFailed + COREFL_FAIL_NONFINITE;
      return;
    }
    const double rem = P.t_end[slot] - P.time[slot];
    const int last = m >= rem ? 1 : 0;
    P.dt[slot] = last ? rem : m;
    P.last[slot] = last;
  }
}

__global__ __launch_bounds__(kBlock) void k_tube_dtc(TubePool P, const int *__restrict__ map, double *dtc, long n) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  const int s = map[t];
  dtc[t] = (s >= 0 && P.status[s] == 1) ? P.dt[s] : 0.0;
}

__global__ void k_tube_finish(TubePool P, int n_slots) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n_slots || P.status[slot] != 1) return;
  const long steps = ++P.steps[slot];
  if (P.last[slot]) {
    P.time[slot] = P.t_end[slot];
    P.status[slot] = 2;
  } else {
    P.time[slot] += P.dt[slot];
    if (P.max_steps[slot] > 0 && steps >= P.max_steps[slot]) P.status[slot] = kSlotFailed + COREFL_FAIL_CAP;
  }
}

__global__ void k_copy_planes(double *__restrict__ d, const double *__restrict__ s, long stride, long len) {
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= len) return;
  const long o = blockIdx.y * stride + t;
  d[o] = s[o];
}

hipError_t launch_tube_clear(int *map, const int *ranges, int n_ranges, long map_n, hipStream_t s) {
  if (n_ranges <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tube_clear, dim3(n_ranges), dim3(256), 0, s, map, ranges, map_n);
  return hipGetLastError();
}

hipError_t launch_tube_admit(const Zone &z, const TubePool &P, int model, const Tab *tb, const double *rec,
                             int n_rec, int *map, long map_n, hipStream_t s) {
  if (n_rec <= 0) return hipSuccess;
  (void)tb;
  COREFL_DISPATCH_1T(model, hipLaunchKernelGGL((k_tube_admit<MV>), dim3(n_rec), dim3(256), 0, s, z, P, rec, map,
                                               map_n));
  return hipGetLastError();
}

hipError_t launch_tube_cv(const Zone &z, int model, const Tab *tb, const double *rec, int n_rec, hipStream_t s) {
  if (n_rec <= 0) return hipSuccess;
  COREFL_DISPATCH_1T(model, hipLaunchKernelGGL((k_tube_cv<MV>), dim3(n_rec), dim3(256), 0, s, z, tb, rec));
  return hipGetLastError();
}

hipError_t launch_tube_bc(const Zone &z, const TubePool &P, int model, const Tab *tb, int n_slots, hipStream_t s) {
  if (n_slots <= 0) return hipSuccess;
  const long n = 2L * (z.ngg + 1) * n_slots;
  COREFL_DISPATCH_1T(model, hipLaunchKernelGGL((k_tube_bc<MV>), dim3(blocks_for(n, 64)), dim3(64), 0, s, z, P, tb,
                                               n_slots));
  return hipGetLastError();
}

hipError_t launch_tube_dt(const Zone &z, const TubePool &P, int n_slots, hipStream_t s, double fixed_dt) {
  if (n_slots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tube_dt, dim3(n_slots), dim3(kBlock), 0, s, z, P, fixed_dt);
  return hipGetLastError();
}

hipError_t launch_tube_dtc(const Zone &z, const TubePool &P, const int *map, double *dtc, hipStream_t s) {
  const long n = z.mx;
  hipLaunchKernelGGL(k_tube_dtc, dim3(blocks_for(n)), dim3(kBlock), 0, s, P, map, dtc, n);
  return hipGetLastError();
}

hipError_t launch_tube_finish(const TubePool &P, int n_slots, hipStream_t s) {
  if (n_slots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tube_finish, dim3(blocks_for(n_slots, 64)), dim3(64), 0, s, P, n_slots);
  return hipGetLastError();
}

hipError_t launch_properties_row(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  const long n = z.mx + 2L * z.ngg;
  COREFL_DISPATCH_1T(r.model, hipLaunchKernelGGL((k_properties<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                                 tb, 0, n));
  return hipGetLastError();
}

hipError_t launch_cv_from_bv_row(const Zone &z, const Run &r, const Tab *tb, hipStream_t s) {
  const long n = z.mx + 2L * z.ngg;
  COREFL_DISPATCH_1T(r.model, hipLaunchKernelGGL((k_cv_from_bv<MV>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                                 tb, n));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Wu's operator splitting (temporal_scheme 4), src/StrangSplitting.cu:43-566
// ---------------------------------------------------------------------------
// The primitive recovery after a change of cv, in registers: velocities, the
// species clipped at 0 and rescaled to sum to rho, T by the warm-started
// Newton on e(T) (relative 1e-5, at most 1000 iterations) and p
// (src/StrangSplitting.cu:193-227 with compute_temperature_and_pressure,
// src/FieldOperation.cu:4-94; one temperature).
template <int M>
DEV void wu_recover(const Tab *__restrict__ tb, double *cv, double *y, double &T, double &pr, double &u, double &v,
                    double &w) {
  constexpr int NS = kNs(M);
  const double density_inv = 1.0 / cv[0];
  u = cv[1] * density_inv;
  v = cv[2] * density_inv;
  w = cv[3] * density_inv;
  double sum = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    if (cv[5 + l] < 0)
      cv[5 + l] = 0;
    else
      sum += cv[5 + l];
  }
  sum = 1.0 / sum;
  const double ratio = cv[0] * sum;
  double gas_const = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    y[l] = cv[5 + l] * sum;
    cv[5 + l] *= ratio;
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) gas_const += y[l] * tb->gas_const[l];
  const double e = cv[4] / cv[0] - 0.5 * (u * u + v * v + w * w);
  double err = 1, tt = T;
  int iter = 0;
  while (err > 1e-5 && iter++ < 1000) {
    double h[NS], cp[NS];
    thermo_h_cp<M>(tb, tt, h, cp);
    double cp_tot = 0, hh = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp_tot += cp[l] * y[l];
      hh += h[l] * y[l];
    }
    const double t1 = tt - (hh - gas_const * tt - e) / (cp_tot - gas_const);
    err = fabs(1 - t1 / tt);
    tt = t1;
  }
  T = tt;
  pr = cv[0] * tt * gas_const;
}

// One wu_zeroDReaction sub-step of length h (src/StrangSplitting.cu:43-228):
// momentum and energy advance by the frozen transport part Cn, the partial
// densities by the chemical source plus theirs; method 3 takes three SSP-RK3
// stages with the rates at each stage's T, 1 divides the source by I - h J
// (the mass-action Jacobian), 0 is explicit.  time_scale: min |rho Y / omega_d|.
template <int M, int METHOD>
DEV void wu_substep(const Tab *__restrict__ tb, double *cv, const double *cn, double h, double iJac,
                    double *y, double &T, double &pr, double &u, double &v, double &w, double &time_scale) {
  constexpr int NS = kNs(M), NV = kNv(M);
  double qn[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) qn[l] = cv[l];
  double cr[NS], od[NS];
  if constexpr (METHOD == 3) {
    for (int stage = 0; stage < 3; ++stage) {
      const double a = c_rk_a[stage], b = c_rk_b[stage], c = c_rk_c[stage];
      // concentrations from the partial densities: rho = 1 times rho Y
      chem_rates<M, false>(tb, 1.0, T, T, cv + 5, cr, od, nullptr);
#pragma unroll
      for (int l = 0; l < 5; ++l) cv[l] = a * qn[l] + b * cv[l] + c * h * iJac * cn[l];
      double ts = 1e+6;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const double conv = 1e+3 / tb->imw[l];
        const double odl = od[l] * conv, om = cr[l] * conv - odl;
        if (cv[5 + l] > 1e-25 && fabs(odl) > 1e-25) ts = fmin(ts, fabs(cv[5 + l] / odl));
        cv[5 + l] = a * qn[5 + l] + b * cv[5 + l] + c * h * (om + cn[5 + l] * iJac);
      }
      time_scale = ts;
      wu_recover<M>(tb, cv, y, T, pr, u, v, w);
    }
    return;
  }
  double om[NS];
  if constexpr (METHOD == 1) {
    double zz[NS * NS];
    chem_rates<M, true>(tb, 1.0, T, T, cv + 5, cr, od, zz);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double conv = 1e+3 / tb->imw[l];
      om[l] = cr[l] * conv - od[l] * conv;
    }
    // (I - h J) x = omega, J[m][n] = 1e3 / imw_m * zz[m][n] / (rho Y_n) where rho Y_n > 1e-25
    // (rho Y_n as the reference forms it back from the concentration)
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double ry = cv[5 + n] * tb->imw[n] * 1e-3 * (1e+3 / tb->imw[n]);
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        double jmn = 0;
        if (ry > 1e-25) jmn = zz[m * NS + n] / ry;
        zz[m * NS + n] = (m == n ? 1.0 : 0.0) - h * (jmn * 1e+3 / tb->imw[m]);
      }
    }
    int ipiv[NS];
    lu_factor<NS>(zz, ipiv);
    lu_apply<NS>(zz, ipiv, om);
  } else {
    chem_rates<M, false>(tb, 1.0, T, T, cv + 5, cr, od, nullptr);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double conv = 1e+3 / tb->imw[l];
      om[l] = cr[l] * conv - od[l] * conv;
    }
  }
#pragma unroll
  for (int l = 0; l < 5; ++l) cv[l] += cn[l] * h * iJac;
  double ts = 1e+6;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double odl = od[l] * (1e+3 / tb->imw[l]);
    if (cv[5 + l] > 1e-25 && fabs(odl) > 1e-25) ts = fmin(ts, fabs(cv[5 + l] / odl));
    cv[5 + l] += (om[l] + cn[5 + l] * iJac) * h;
  }
  time_scale = ts;
  wu_recover<M>(tb, cv, y, T, pr, u, v, w);
}

// wu_reaction_step over the interior (src/StrangSplitting.cu:377-425): the
// point's reaction time scale (0.2 dt when unset) splits dt into n_sub
// sub-steps of dt / (n_sub + 1) and a last one of what is left (skipped
// below 1e-12).  dtc: a dt per point (batched tubes); a dt of 0 leaves the
// point as it is.
// (one instantiation per sub-step method: the point-implicit LU's registers do not weigh on the others)
template <int M, int METHOD>
__global__ __launch_bounds__(kBlock) void k_wu_reaction(Zone z, const Tab *__restrict__ tb, double dt,
                                                        const double *__restrict__ dtc, long n) {
  constexpr int NS = kNs(M), NV = kNv(M);
  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  if (dtc) dt = dtc[t];
  int i, j, k;
  interior_ijk(z, t, i, j, k);
  const long p = z.ix(i, j, k), sz = z.sz, dsz = z.dsz;
  double ts = z.rts[p];
  if (ts < 1e-15) ts = 0.2 * dt;
  const int n_sub = static_cast<int>(dt / ts);
  const double dtr = dt / (n_sub + 1);
  const double last = dt - dtr * n_sub;
  if (n_sub == 0 && !(fabs(last) > 1e-12)) return;
  double cv[NV], cn[NV], y[NS];
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    cv[l] = z.cv[p + l * sz];
    cn[l] = z.cn[t + l * dsz];
  }
  const double iJac = 1.0 / z.jac[p];
  double T = z.bv[p + 5 * sz], pr = 0, u = 0, v = 0, w = 0, time_scale = 0;
  for (int s = 0; s < n_sub; ++s) wu_substep<M, METHOD>(tb, cv, cn, dtr, iJac, y, T, pr, u, v, w, time_scale);
  if (fabs(last) > 1e-12) wu_substep<M, METHOD>(tb, cv, cn, last, iJac, y, T, pr, u, v, w, time_scale);
#pragma unroll
  for (int l = 0; l < NV; ++l) z.cv[p + l * sz] = cv[l];
  z.bv[p] = cv[0];
  z.bv[p + sz] = u;
  z.bv[p + 2 * sz] = v;
  z.bv[p + 3 * sz] = w;
  z.bv[p + 4 * sz] = pr;
  z.bv[p + 5 * sz] = T;
#pragma unroll
  for (int l = 0; l < NS; ++l) z.sv[p + l * sz] = y[l];
  z.rts[p] = time_scale;
  flag_state(z, pos_finite(cv[0]) && pos_finite(pr) && pos_finite(T) && isfinite(u) && isfinite(v) && isfinite(w));
}

hipError_t launch_wu_reaction(const Zone &z, const Run &r, const Tab *tb, double dt, int method, hipStream_t s,
                              const double *dtc) {
  const long n = interior_points(z);
  if (method == 1) {
    COREFL_DISPATCH_1T(r.model, hipLaunchKernelGGL((k_wu_reaction<MV, 1>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                                   tb, dt, dtc, n));
  } else if (method == 3) {
    COREFL_DISPATCH_1T(r.model, hipLaunchKernelGGL((k_wu_reaction<MV, 3>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                                   tb, dt, dtc, n));
  } else {
    COREFL_DISPATCH_1T(r.model, hipLaunchKernelGGL((k_wu_reaction<MV, 0>), dim3(blocks_for(n)), dim3(kBlock), 0, s, z,
                                                   tb, dt, dtc, n));
  }
  return hipGetLastError();
}

hipError_t launch_update_rk_wu(const Zone &z, const Run &r, const Tab *tb, double dt, int rk, hipStream_t s,
                               const double *dtc) {
  const long n = interior_points(z);
  COREFL_DISPATCH_1T(r.model, hipLaunchKernelGGL((k_update_rk<MV, true>), dim3(blocks_for(n)), dim3(kBlock), 0, s,
                                                 z, tb, dt, nullptr, dtc, rk, n));
  return hipGetLastError();
}

hipError_t launch_copy_planes(double *dst, const double *src, int planes, long stride, long len, hipStream_t s) {
  if (planes <= 0 || len <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_planes, dim3(blocks_for(len), planes), dim3(kBlock), 0, s, dst, src, stride, len);
  return hipGetLastError();
}

}  // namespace corefl
