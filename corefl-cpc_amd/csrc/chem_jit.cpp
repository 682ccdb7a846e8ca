// Mechanism-specialised finite-rate chemistry kernels, generated when a context
// is created and compiled with hipRTC for the device's architecture.  The
// kernels perform the table-driven kernels' arithmetic in the same order
// (k_chemistry in kernels.hip, restating src/FiniteRateChem.cu:1406-1681, and
// Wu's explicit reaction step k_wu_reaction<M, 0>, src/StrangSplitting.cu:43-425),
// with every constant of the mechanism written as an exact hexadecimal literal
// and the loops over reactions and stoichiometric lists unrolled, so the
// per-reaction table reads and branches become straight-line FP64 code.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "corefl_internal.h"

namespace corefl {
namespace {

std::string lit(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "(%a)", v);
  return buf;
}

// NASA-7 G/RT of species i at T (thermo_gibbs in kernels.hip): the in-range
// branch with the low or high coefficient set, linear extrapolation outside.
void emit_gibbs(std::ostringstream &o, const Tab &tb, int i) {
  const double *lo = tb.low_coeff[i], *hi = tb.high_coeff[i];
  auto g_expr = [&](const double *c, const char *t, const char *t2, const char *t3, const char *t4, const char *ti,
                    const char *lt) {
    std::ostringstream e;
    e << lit(c[0]) << " * (1.0 - " << lt << ") - 0.5 * " << lit(c[1]) << " * " << t << " - " << lit(c[2]) << " * "
      << t2 << " / 6.0 - " << lit(c[3]) << " * " << t3 << " / 12.0 - " << lit(c[4]) << " * " << t4 << " * 0.05 + "
      << lit(c[5]) << " * " << ti << " - " << lit(c[6]);
    return e.str();
  };
  o << "  double g" << i << ";\n";
  o << "  if (T < " << lit(tb.t_low[i]) << " || T > " << lit(tb.t_high[i]) << ") {\n";
  o << "    const double tt = T < " << lit(tb.t_low[i]) << " ? " << lit(tb.t_low[i]) << " : " << lit(tb.t_high[i])
    << ";\n";
  o << "    const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);\n";
  o << "    if (T < " << lit(tb.t_low[i]) << ") g" << i << " = " << g_expr(lo, "tt", "tt2", "tt3", "tt4", "tt_inv", "log_tt")
    << ";\n";
  o << "    else g" << i << " = " << g_expr(hi, "tt", "tt2", "tt3", "tt4", "tt_inv", "log_tt") << ";\n";
  o << "  } else if (T < " << lit(tb.t_mid[i]) << ") {\n";
  o << "    g" << i << " = " << g_expr(lo, "T", "t2", "t3", "t4", "t_inv", "log_t") << ";\n";
  o << "  } else {\n";
  o << "    g" << i << " = " << g_expr(hi, "T", "t2", "t3", "t4", "t_inv", "log_t") << ";\n";
  o << "  }\n";
}

std::string ipow_expr(int l, int e) {
  if (e == 0) return "1.0";
  if (e == 1) return "c" + std::to_string(l);
  if (e == 2) return "(c" + std::to_string(l) + " * c" + std::to_string(l) + ")";
  return "pow(c" + std::to_string(l) + ", " + lit(static_cast<double>(e)) + ")";
}

std::string arrhenius(double A, double b, double Ea) {
  return lit(A) + " * exp(" + lit(b) + " * ln_t - " + lit(Ea) + " * inv_t)";
}

// The rates of every species from the concentrations c_l at T (chem_rates
// in kernels.hip, one temperature): creation cr_l and destruction od_l, with
// the mechanism's constants as literals and the reaction loop unrolled.
void emit_rates(std::ostringstream &o, const Tab &tb, int ns) {
  o << "__device__ __forceinline__ void corefl_rates(const double *cc, double T, double *cr, double *od) {\n";
  for (int l = 0; l < ns; ++l) o << "  const double c" << l << " = cc[" << l << "];\n";
  o << "  const double t2 = T * T, t3 = t2 * T, t4 = t3 * T, t_inv = 1 / T, log_t = log(T);\n";
  for (int l = 0; l < ns; ++l) emit_gibbs(o, tb, l);
  o << "  const double ln_t = log(T), inv_t = 1.0 / T;\n";
  o << "  const double ln_temp_t = log(" << lit(101325.0 / kRu * 1e-3) << ") - ln_t;\n";
  for (int l = 0; l < ns; ++l) o << "  double cr" << l << " = 0, od" << l << " = 0;\n";
  for (int r = 0; r < tb.n_reac; ++r) {
    const Reac &R = tb.reac[r];
    const int type = tb.reac_type[r];
    o << "  {  // reaction " << r << "\n";
    o << "    double kf = " << arrhenius(tb.A[r], tb.b[r], tb.Ea[r]) << ";\n";
    if (type > 3) {
      // zero efficiencies add exact zeros to the table-driven kernel's sum: skipped
      o << "    double cc = 0;\n";
      for (int l = 0; l < ns; ++l)
        if (tb.third_body[r][l] != 0.0) o << "    cc += c" << l << " * " << lit(tb.third_body[r][l]) << ";\n";
    }
    if (type == 3) {
      o << "    kf += " << arrhenius(tb.A2[r], tb.b2[r], tb.Ea2[r]) << ";\n";
    } else if (type == 4) {
      o << "    kf *= cc;\n";
    } else if (type > 4) {
      o << "    {\n      const double kf_low = " << arrhenius(tb.A2[r], tb.b2[r], tb.Ea2[r]) << ";\n";
      o << "      const double kf_high = kf;\n";
      o << "      if (kf_high < 1e-25 && kf_low < 1e-25) {\n        kf = 0;\n      } else {\n";
      o << "        const double reduced_pressure = kf_low * cc / kf_high;\n        double F = 1.0;\n";
      if (type > 5) {
        o << "        double f_cent = (1 - " << lit(tb.troe_alpha[r]) << ") * exp(-T / " << lit(tb.troe_t3[r])
          << ") + " << lit(tb.troe_alpha[r]) << " * exp(-T / " << lit(tb.troe_t1[r]) << ");\n";
        if (type == 7) o << "        f_cent += exp(-" << lit(tb.troe_t2[r]) << " * inv_t);\n";
        o << "        const double logFc = log(f_cent) * 0.4342944819032518;\n"
             "        const double cT = -0.4 - 0.67 * logFc;\n"
             "        const double nn = 0.75 - 1.27 * logFc;\n"
             "        const double logPr = log(reduced_pressure) * 0.4342944819032518;\n"
             "        const double tempo = (logPr + cT) / (nn - 0.14 * (logPr + cT));\n"
             "        F = exp(2.302585092994046 * (logFc / (1.0 + tempo * tempo)));\n";
      }
      o << "        kf = kf_high * reduced_pressure / (1.0 + reduced_pressure) * F;\n      }\n    }\n";
    }
    o << "    double kb = 0;\n";
    if (type != 0) {
      if (tb.rev_type[r] == 1) {
        o << "    kb = " << arrhenius(tb.A2[r], tb.b2[r], tb.Ea2[r]) << ";\n";
        if (type == 4) o << "    kb *= cc;\n";
      }
      if (type != 2) {
        o << "    double d_gibbs = 0;\n";
        for (int m = 0; m < R.nnu; ++m) o << "    d_gibbs += g" << R.nu_idx[m] << " * " << lit(R.nu[m]) << ";\n";
        o << "    kb = kf * exp(d_gibbs - " << lit(static_cast<double>(tb.order[r])) << " * ln_temp_t);\n";
      }
    }
    o << "    double q1 = 1.0, q2 = 1.0;\n";
    for (int m = 0; m < R.nf; ++m) o << "    q1 *= " << ipow_expr(R.f_idx[m], R.f_ord[m]) << ";\n";
    o << "    q1 *= kf;\n";
    if (type != 0) {
      for (int m = 0; m < R.nb; ++m) o << "    q2 *= " << ipow_expr(R.b_idx[m], R.b_ord[m]) << ";\n";
      o << "    q2 *= kb;\n";
    } else {
      o << "    q2 = 0.0;\n";
    }
    for (int m = 0; m < R.npart; ++m) {
      const int l = R.part_idx[m];
      o << "    cr" << l << " += q2 * " << lit(R.part_f[m]) << " + q1 * " << lit(R.part_b[m]) << ";\n";
      o << "    od" << l << " += q1 * " << lit(R.part_f[m]) << " + q2 * " << lit(R.part_b[m]) << ";\n";
    }
    o << "  }\n";
  }
  for (int l = 0; l < ns; ++l) o << "  cr[" << l << "] = cr" << l << ";\n  od[" << l << "] = od" << l << ";\n";
  o << "}\n\n";
}

// NASA-7 h_l and cp_l at t (thermo_h_cp in kernels.hip): h and cp of the low
// or high set in range, linear extrapolation of h with the bound's cp outside.
void emit_h_cp(std::ostringstream &o, const Tab &tb, int ns) {
  auto h_expr = [&](const double *c, const char *t, const char *t2, const char *t3, const char *t4, const char *t5) {
    std::ostringstream e;
    e << lit(c[0]) << " * " << t << " + 0.5 * " << lit(c[1]) << " * " << t2 << " + " << lit(c[2]) << " * " << t3
      << " / 3 + 0.25 * " << lit(c[3]) << " * " << t4 << " + 0.2 * " << lit(c[4]) << " * " << t5 << " + "
      << lit(c[5]);
    return e.str();
  };
  auto cp_expr = [&](const double *c, const char *t, const char *t2, const char *t3, const char *t4) {
    std::ostringstream e;
    e << lit(c[0]) << " + " << lit(c[1]) << " * " << t << " + " << lit(c[2]) << " * " << t2 << " + " << lit(c[3])
      << " * " << t3 << " + " << lit(c[4]) << " * " << t4;
    return e.str();
  };
  o << "__device__ __forceinline__ void corefl_h_cp(double t, double *h, double *cp) {\n";
  o << "  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;\n";
  for (int i = 0; i < ns; ++i) {
    const double *lo = tb.low_coeff[i], *hi = tb.high_coeff[i];
    o << "  {\n    double vh, vc;\n";
    o << "    if (t < " << lit(tb.t_low[i]) << " || t > " << lit(tb.t_high[i]) << ") {\n";
    for (int side = 0; side < 2; ++side) {
      const double *c = side == 0 ? lo : hi;
      const double b = side == 0 ? tb.t_low[i] : tb.t_high[i];
      o << (side == 0 ? "      if (t < " + lit(tb.t_low[i]) + ") {\n" : "      else {\n");
      o << "        const double tt = " << lit(b) << ", tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, "
        << "tt5 = tt4 * tt;\n";
      o << "        vh = " << h_expr(c, "tt", "tt2", "tt3", "tt4", "tt5") << ";\n";
      o << "        vc = " << cp_expr(c, "tt", "tt2", "tt3", "tt4") << ";\n";
      o << "        vh += vc * (t - tt);\n      }\n";
    }
    o << "    } else if (t < " << lit(tb.t_mid[i]) << ") {\n";
    o << "      vh = " << h_expr(lo, "t", "t2", "t3", "t4", "t5") << ";\n";
    o << "      vc = " << cp_expr(lo, "t", "t2", "t3", "t4") << ";\n";
    o << "    } else {\n";
    o << "      vh = " << h_expr(hi, "t", "t2", "t3", "t4", "t5") << ";\n";
    o << "      vc = " << cp_expr(hi, "t", "t2", "t3", "t4") << ";\n";
    o << "    }\n";
    o << "    h[" << i << "] = vh * " << lit(tb.gas_const[i]) << ";\n";
    o << "    cp[" << i << "] = vc * " << lit(tb.gas_const[i]) << ";\n  }\n";
  }
  o << "}\n\n";
}

// The point's source terms (k_chemistry in kernels.hip): dq += jac omega and
// the reaction time scale.
void emit_chem_kernel(std::ostringstream &o, const Tab &tb, int ns) {
  o << "extern \"C\" __global__ __launch_bounds__(256) void corefl_chem(const double *__restrict__ bv,\n"
       "    const double *__restrict__ sv, const double *__restrict__ jacobian, double *__restrict__ dq,\n"
       "    double *__restrict__ rts, long sz, long sx, long sxy, long off, int mx, int my, long n, long dsz) {\n"
       "  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;\n"
       "  if (t >= n) return;\n"
       "  const int i = static_cast<int>(t % mx);\n"
       "  const long r_ = t / mx;\n"
       "  const int j = static_cast<int>(r_ % my);\n"
       "  const int k = static_cast<int>(r_ / my);\n"
       "  const long p = k * sxy + j * sx + i + off;\n"
       "  const double density = bv[p];\n"
       "  const double T = bv[p + 5 * sz];\n";
  o << "  double c[" << ns << "], cr[" << ns << "], od[" << ns << "];\n";
  for (int l = 0; l < ns; ++l) {
    o << "  const double y" << l << " = sv[p + " << l << " * sz];\n";
    o << "  c[" << l << "] = density * y" << l << " * " << lit(tb.imw[l]) << " * 1e-3;\n";
  }
  o << "  corefl_rates(c, T, cr, od);\n";
  o << "  const double jac = jacobian[p];\n  double time_scale = 1e+6;\n";
  for (int l = 0; l < ns; ++l) {
    o << "  {\n    const double conv = 1e+3 / " << lit(tb.imw[l]) << ";\n";
    o << "    const double crl = cr[" << l << "] * conv, odl = od[" << l << "] * conv;\n";
    o << "    const double om = crl - odl;\n";
    o << "    dq[t + " << (l + 5) << " * dsz] += jac * om;\n";
    o << "    if (y" << l << " > 1e-25 && fabs(odl) > 1e-25) time_scale = fmin(time_scale, fabs(density * y" << l
      << " / odl));\n  }\n";
  }
  o << "  rts[p] = time_scale;\n}\n\n";
}

// Wu's reaction step with explicit sub-steps (k_wu_reaction<M, 0> with
// wu_substep and wu_recover in kernels.hip, src/StrangSplitting.cu:43-228,
// 377-425): the same sub-step split, sub-step arithmetic and recovery.
void emit_wu_kernel(std::ostringstream &o, const Tab &tb, int ns) {
  const int nv = ns + 5;
  o << "extern \"C\" __global__ __launch_bounds__(256) void corefl_wu(double *__restrict__ cvp,\n"
       "    const double *__restrict__ cnp, double *__restrict__ bv, double *__restrict__ sv,\n"
       "    double *__restrict__ rts, const double *__restrict__ jacobian, const double *__restrict__ dtc,\n"
       "    double dt, int *bad, long sz, long sx, long sxy, long off, int mx, int my, long n, long dsz) {\n"
       "  const long t = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;\n"
       "  if (t >= n) return;\n"
       "  if (dtc) dt = dtc[t];\n"
       "  const int i = static_cast<int>(t % mx);\n"
       "  const long r_ = t / mx;\n"
       "  const int j = static_cast<int>(r_ % my);\n"
       "  const int k = static_cast<int>(r_ / my);\n"
       "  const long p = k * sxy + j * sx + i + off;\n"
       "  double ts0 = rts[p];\n"
       "  if (ts0 < 1e-15) ts0 = 0.2 * dt;\n"
       "  const int n_sub = static_cast<int>(dt / ts0);\n"
       "  const double dtr = dt / (n_sub + 1);\n"
       "  const double last = dt - dtr * n_sub;\n"
       "  if (n_sub == 0 && !(fabs(last) > 1e-12)) return;\n";
  o << "  double cv[" << nv << "], y[" << ns << "];\n";
  o << "  for (int l = 0; l < " << nv << "; ++l) cv[l] = cvp[p + l * sz];\n";
  o << "  const double iJac = 1.0 / jacobian[p];\n";
  o << "  double T = bv[p + 5 * sz], pr = 0, u = 0, v = 0, w = 0, time_scale = 0;\n";
  // the n_sub equal sub-steps, then the remainder when it is above 1e-12
  o << "  const int total = n_sub + (fabs(last) > 1e-12 ? 1 : 0);\n";
  o << "  for (int s = 0; s < total; ++s) {\n";
  // Cn is read in every sub-step through a pointer the loop re-derives: held in registers across the
  // loop it would cost the second wave per SIMD
  o << "    const double *cn = cnp + t;\n";
  o << "    asm volatile(\"\" : \"+v\"(cn));\n";
  o << "    const double h = s < n_sub ? dtr : last;\n";
  o << "    double c[" << ns << "], cr[" << ns << "], od[" << ns << "];\n";
  for (int l = 0; l < ns; ++l) o << "    c[" << l << "] = cv[" << (5 + l) << "] * " << lit(tb.imw[l]) << " * 1e-3;\n";
  o << "    corefl_rates(c, T, cr, od);\n";
  o << "    for (int l = 0; l < 5; ++l) cv[l] += cn[l * dsz] * h * iJac;\n";
  o << "    double ts = 1e+6;\n";
  for (int l = 0; l < ns; ++l) {
    const int q = 5 + l;
    o << "    {\n      const double conv = 1e+3 / " << lit(tb.imw[l]) << ";\n";
    o << "      const double om = cr[" << l << "] * conv - od[" << l << "] * conv;\n";
    o << "      const double odl = od[" << l << "] * conv;\n";
    o << "      if (cv[" << q << "] > 1e-25 && fabs(odl) > 1e-25) ts = fmin(ts, fabs(cv[" << q << "] / odl));\n";
    o << "      cv[" << q << "] += (om + cn[" << q << " * dsz] * iJac) * h;\n    }\n";
  }
  o << "    time_scale = ts;\n";
  // recovery: velocities, clipped and rescaled species, T by Newton on e(T), p
  o << "    const double density_inv = 1.0 / cv[0];\n"
       "    u = cv[1] * density_inv;\n    v = cv[2] * density_inv;\n    w = cv[3] * density_inv;\n"
       "    double sum = 0;\n";
  o << "    for (int l = 5; l < " << nv << "; ++l) {\n      if (cv[l] < 0)\n        cv[l] = 0;\n"
       "      else\n        sum += cv[l];\n    }\n";
  o << "    sum = 1.0 / sum;\n    const double ratio = cv[0] * sum;\n";
  o << "    for (int l = 0; l < " << ns << "; ++l) {\n      y[l] = cv[5 + l] * sum;\n      cv[5 + l] *= ratio;\n"
       "    }\n";
  o << "    double gas_const = 0;\n";
  for (int l = 0; l < ns; ++l) o << "    gas_const += y[" << l << "] * " << lit(tb.gas_const[l]) << ";\n";
  o << "    const double e = cv[4] / cv[0] - 0.5 * (u * u + v * v + w * w);\n"
       "    double err = 1, tt = T;\n    int iter = 0;\n"
       "    while (err > 1e-5 && iter++ < 1000) {\n"
       "      double hs[" << ns << "], cps[" << ns << "];\n"
       "      corefl_h_cp(tt, hs, cps);\n"
       "      double cp_tot = 0, hh = 0;\n"
       "      for (int l = 0; l < " << ns << "; ++l) {\n        cp_tot += cps[l] * y[l];\n"
       "        hh += hs[l] * y[l];\n      }\n"
       "      const double t1 = tt - (hh - gas_const * tt - e) / (cp_tot - gas_const);\n"
       "      err = fabs(1 - t1 / tt);\n      tt = t1;\n    }\n"
       "    T = tt;\n    pr = cv[0] * tt * gas_const;\n";
  o << "  }\n";
  o << "  for (int l = 0; l < " << nv << "; ++l) cvp[p + l * sz] = cv[l];\n";
  o << "  bv[p] = cv[0];\n  bv[p + sz] = u;\n  bv[p + 2 * sz] = v;\n  bv[p + 3 * sz] = w;\n"
       "  bv[p + 4 * sz] = pr;\n  bv[p + 5 * sz] = T;\n";
  o << "  for (int l = 0; l < " << ns << "; ++l) sv[p + l * sz] = y[l];\n";
  o << "  rts[p] = time_scale;\n";
  o << "  const double inf = __builtin_inf();\n";
  o << "  const bool ok = cv[0] > 0.0 && cv[0] < inf && pr > 0.0 && pr < inf && T > 0.0 && T < inf &&\n"
       "                  __builtin_isfinite(u) && __builtin_isfinite(v) && __builtin_isfinite(w);\n";
  o << "  if (!ok && bad) *bad = 1;\n}\n";
}

std::string generate(const Tab &tb, int ns) {
  std::ostringstream o;
  emit_rates(o, tb, ns);
  emit_h_cp(o, tb, ns);
  emit_chem_kernel(o, tb, ns);
  emit_wu_kernel(o, tb, ns);
  return o.str();
}

std::mutex g_mu;
std::map<std::tuple<int, std::string, std::string>, ChemJit> g_cache;  // (device, arch, source) -> kernel

// Compiles generated source to a code object for arch (the device's gcnArchName, e.g.
// "gfx950:sramecc+:xnack-", so the object matches the device's target features).
Status compile(const std::string &src, const std::string &arch, std::vector<char> &code, std::string &err) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "corefl_chem.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return Status::Hip;
  }
  // the same code generation switches as the library's own kernels; machine-level loop-invariant code
  // motion off: it would hoist the mechanism's literals out of the Wu sub-step loop into registers held
  // across it (one wave per SIMD instead of two); the source-term kernel has no loop
  const std::string arch_opt = "--offload-arch=" + arch;
  const char *opts[] = {arch_opt.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops", "-mllvm", "-disable-machine-licm"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 9, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    err = std::string("hipRTC: ") + hiprtcGetErrorString(rc) + "\n" + log;
    hiprtcDestroyProgram(&prog);
    return Status::Hip;
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  code.resize(code_size);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return Status::Ok;
}

}  // namespace

Status chem_jit_build(const Tab &tb, int ns, ChemJit &out, std::string &err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    err = "no current device";
    return Status::Hip;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    err = "hipGetDeviceProperties failed";
    return Status::Hip;
  }
  const std::string arch = prop.gcnArchName;
  const std::string src = generate(tb, ns);
  std::lock_guard<std::mutex> lock(g_mu);
  const auto key = std::make_tuple(dev, arch, src);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    out = it->second;
    return Status::Ok;
  }
  std::vector<char> code;
  const Status st = compile(src, arch, code, err);
  if (st != Status::Ok) return st;
  ChemJit j{};
  hipError_t e = hipModuleLoadData(&j.mod, code.data());
  if (e == hipSuccess) e = hipModuleGetFunction(&j.fn, j.mod, "corefl_chem");
  if (e == hipSuccess) e = hipModuleGetFunction(&j.wu, j.mod, "corefl_wu");
  if (e != hipSuccess) {
    err = std::string("loading the chemistry kernel: ") + hipGetErrorString(e);
    return Status::Hip;
  }
  g_cache[key] = j;
  out = j;
  return Status::Ok;
}

std::string chem_jit_source(const Tab &tb, int ns) { return generate(tb, ns); }

hipError_t chem_jit_launch(const ChemJit &j, const Zone &z, hipStream_t s) {
  const double *bv = z.bv, *sv = z.sv, *jac = z.jac;
  double *dq = z.dq, *rts = z.rts;
  long sz = z.sz, sx = z.sx, sxy = z.sxy, off = z.off, dsz = z.dsz;
  int mx = z.mx, my = z.my;
  long n = static_cast<long>(z.mx) * z.my * z.mz;
  void *args[] = {&bv, &sv, &jac, &dq, &rts, &sz, &sx, &sxy, &off, &mx, &my, &n, &dsz};
  const unsigned nb = static_cast<unsigned>((n + 255) / 256);
  return hipModuleLaunchKernel(j.fn, nb, 1, 1, 256, 1, 1, 0, s, args, nullptr);
}

hipError_t chem_jit_launch_wu(const ChemJit &j, const Zone &z, double dt, const double *dtc, hipStream_t s) {
  double *cv = z.cv, *bv = z.bv, *sv = z.sv, *rts = z.rts;
  const double *cn = z.cn, *jac = z.jac;
  int *bad = z.bad;
  long sz = z.sz, sx = z.sx, sxy = z.sxy, off = z.off, dsz = z.dsz;
  int mx = z.mx, my = z.my;
  long n = static_cast<long>(z.mx) * z.my * z.mz;
  void *args[] = {&cv, &cn, &bv, &sv, &rts, &jac, &dtc, &dt, &bad, &sz, &sx, &sxy, &off, &mx, &my, &n, &dsz};
  const unsigned nb = static_cast<unsigned>((n + 255) / 256);
  return hipModuleLaunchKernel(j.wu, nb, 1, 1, 256, 1, 1, 0, s, args, nullptr);
}

}  // namespace corefl

// Generates and compiles the mechanism's chemistry kernel without loading it
// (no device needed).  Returns 0, or COREFL_ERR_HIP with the compiler log.
extern "C" int corefl_chem_kernel_check(const corefl_chem_tables *tables, char *log, size_t log_cap) {
  using namespace corefl;
  if (!tables || tables->n_spec < 1 || tables->n_spec > COREFL_MAX_SPEC || tables->n_reac < 0 ||
      tables->n_reac > COREFL_MAX_REAC)
    return static_cast<int>(Status::Arg);
  auto tb = std::make_unique<Tab>();
  derive_tables(*tables, *tb);
  std::vector<char> code;
  std::string err;
  // the library's own target when no device is current
  std::string arch = "gfx950";
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) arch = prop.gcnArchName;
  else (void)hipGetLastError();
  const Status st = compile(generate(*tb, tables->n_spec), arch, code, err);
  if (log && log_cap) {
    std::snprintf(log, log_cap, "%s", st == Status::Ok ? "" : err.c_str());
  }
  return static_cast<int>(st);
}
