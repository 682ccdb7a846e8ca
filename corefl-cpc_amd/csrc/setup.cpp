// Reader for the reference's setup.txt format (src/Parameter.cpp:38-184):
// lines "type name = value" with type int/real/bool/string, "array <type> name
// { ... }" and "struct name { ... }" blocks, comments starting with // or !.
// Fills corefl_run_desc with the keys the per-stage path reads, applying the
// reference defaults (src/Parameter.cpp:358-546) and derived counts
// (src/Parameter.cpp:186-356, 571-687), and the reference state
// (Inflow constructor, src/BoundCond.cpp:215-357).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <variant>
#include <vector>

#include "corefl_internal.h"

namespace corefl {

bool read_setup(const std::string &path, Setup &s, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::string line;
  while (std::getline(f, line)) {
    if (line.rfind("//", 0) == 0 || line.rfind("!", 0) == 0 || line.empty()) continue;
    std::istringstream ss(line);
    std::string type, key, eq;
    ss >> type >> key >> eq;
    try {
      if (type == "int") {
        int v = 0;
        ss >> v;
        s.ints[key] = v;
      } else if (type == "real") {
        double v = 0;
        ss >> v;
        s.reals[key] = v;
      } else if (type == "bool") {
        int v = 0;
        ss >> v;
        s.bools[key] = v != 0;
      } else if (type == "string") {
        std::string v;
        ss >> v;
        s.strings[key] = v;
      } else if (type == "array") {
        // "array <type> <name> {" followed by values until "}"
        const std::string name = eq;
        std::string tok;
        ss >> tok;  // "{"
        std::vector<std::string> vals;
        bool done = false;
        auto eat = [&](std::istringstream &in) {
          std::string t;
          while (in >> t) {
            if (t == "}") return true;
            if (t == "//") break;
            if (t == "{") continue;
            vals.push_back(t);
          }
          return false;
        };
        done = eat(ss);
        while (!done && std::getline(f, line)) {
          std::istringstream ls(line);
          done = eat(ls);
        }
        if (key == "string") s.string_arrays[name] = vals;
      } else if (type == "struct") {
        auto &st = s.structs[key];
        while (std::getline(f, line)) {
          std::istringstream ls(line);
          std::string t, name, e2;
          ls >> t;
          if (t == "}") break;
          if (t == "string") {
            std::string v;
            ls >> name >> e2 >> v;
            st.emplace(name, v);
          } else if (t == "int") {
            int v = 0;
            ls >> name >> e2 >> v;
            st.emplace(name, v);
          } else if (t == "real") {
            double v = 0;
            ls >> name >> e2 >> v;
            st.emplace(name, v);
          }
        }
      }
    } catch (const std::exception &) {
      err = "malformed line: " + line;
      return false;
    }
  }
  return true;
}

double struct_real(const SetupStruct &st, const std::string &k, double def) {
  auto it = st.find(k);
  if (it == st.end()) return def;
  if (auto p = std::get_if<double>(&it->second)) return *p;
  if (auto p = std::get_if<int>(&it->second)) return *p;
  return def;
}

namespace {

// NASA-7 cp of species l at T (J/(kg K)), as src/Thermo.cu:551-572.
double host_cp(const corefl_chem_tables &t, int l, double T) {
  if (t.nasa7or9 == 9) {
    const int n = t.n_range[l];
    const double tt = T < t.t_cut[l][0] ? t.t_cut[l][0] : (T > t.t_cut[l][n] ? t.t_cut[l][n] : T);
    int j = 0;
    while (j < n - 1 && t.t_cut[l][j + 1] < tt) ++j;
    const double *c = t.poly9[l][j];
    return (c[0] / (tt * tt) + c[1] / tt + c[2] + c[3] * tt + c[4] * tt * tt + c[5] * tt * tt * tt +
            c[6] * tt * tt * tt * tt) *
           t.gas_const[l];
  }
  const double tt = T < t.t_low[l] ? t.t_low[l] : (T > t.t_high[l] ? t.t_high[l] : T);
  const double *c = tt < t.t_mid[l] ? t.low_coeff[l] : t.high_coeff[l];
  return (c[0] + c[1] * tt + c[2] * tt * tt + c[3] * tt * tt * tt + c[4] * tt * tt * tt * tt) * t.gas_const[l];
}

}  // namespace

// Species names in mechanism order (the tables carry no names).
std::string species_name(const corefl_chem_tables &, const std::string &dir, const Setup &s, int l) {
  const std::string mech = dir + get_or(s.strings, std::string("mechanism_file"), std::string("chemistry/mech.inp"));
  std::ifstream f(mech);
  std::vector<std::string> toks;
  std::string line;
  while (std::getline(f, line)) {
    auto p = line.find('!');
    if (p != std::string::npos) line.resize(p);
    std::istringstream ss(line);
    std::string t;
    while (ss >> t) toks.push_back(t);
  }
  size_t i = 0;
  while (i < toks.size() && toks[i] != "SPECIES" && toks[i] != "SPEC") ++i;
  ++i;
  int n = 0;
  for (; i < toks.size(); ++i) {
    if (toks[i] == "END") continue;
    if (toks[i] == "REACTIONS" || toks[i] == "REAC" || toks[i] == "THERMO") break;
    if (n++ == l) return toks[i];
  }
  return std::string();
}

}  // namespace corefl

using namespace corefl;

extern "C" int corefl_setup_load(const char *case_dir, corefl_run_desc *run, corefl_chem_tables *tables) {
  if (!case_dir || !run) return static_cast<int>(Status::Arg);
  const std::string dir = std::string(case_dir) + "/input/";
  Setup s;
  std::string err;
  if (!read_setup(dir + "setup.txt", s, err)) {
    std::fprintf(stderr, "corefl_setup_load: %s\n", err.c_str());
    return static_cast<int>(Status::Io);
  }
  std::memset(run, 0, sizeof(*run));
  const int species = get_or(s.ints, std::string("species"), 0);
  run->mixture = species != 0;
  run->inviscid_scheme = get_or(s.ints, std::string("inviscid_scheme"), 3);
  run->viscous_order = get_or(s.ints, std::string("viscous_order"), 2);
  run->reaction = species ? get_or(s.ints, std::string("reaction"), 0) : 0;
  run->gradp_in_diffusion = get_or(s.bools, std::string("gradPInDiffusionFlux"), false);
  run->fixed_time_step = get_or(s.bools, std::string("fixed_time_step"), false);
  run->dt = get_or(s.reals, std::string("dt"), 1e-8);
  run->cfl = get_or(s.reals, std::string("cfl"), 0.8);
  run->prandtl = get_or(s.reals, std::string("prandtl_number"), 0.72);
  run->reconstruction = get_or(s.ints, std::string("reconstruction"), 2);
  run->limiter = get_or(s.ints, std::string("limiter"), 0);
  run->entropy_fix_factor = get_or(s.reals, std::string("entropy_fix_factor"), 0.125);
  run->limit_flow = get_or(s.bools, std::string("limit_flow"), false);
  // WENO options, defaults src/Parameter.cpp:404-406, 415
  run->positive_preserving = get_or(s.bools, std::string("positive_preserving"), false);
  run->shock_sensor = get_or(s.ints, std::string("shock_sensor"), 0);
  run->sensor_threshold = get_or(s.reals, std::string("shockSensor_threshold"), 0.0);
  run->sensor_eps = get_or(s.reals, std::string("shockSensor_epsilon"), -1.0);
  // temporal_scheme 3 SSP-RK3 (the default) or 4 Wu's splitting; chem_advance_method for its reaction
  // sub-step (src/Parameter.cpp:380, 453)
  run->temporal_scheme = get_or(s.ints, std::string("temporal_scheme"), 3);
  run->chem_advance_method = get_or(s.ints, std::string("chem_advance_method"), 0);
  if (run->mixture) {
    if (!tables) return static_cast<int>(Status::Arg);
    const int nasa = get_or(s.ints, std::string("therm_nasa_7_9"), 9);
    const std::string mech = dir + get_or(s.strings, std::string("mechanism_file"), std::string("chemistry/mech.inp"));
    const std::string therm = dir + get_or(s.strings, std::string("therm_file"), std::string("chemistry/therm.dat"));
    const std::string tran = dir + get_or(s.strings, std::string("transport_file"), std::string("chemistry/tran.dat"));
    // the two-temperature model (the reference's COREFL_ENABLE_TWO_TEMPERATURE build) is taken on when
    // the setup names its data file; enable_vt_relaxation defaults to on (src/Parameter.cpp:424-426)
    const bool tt = s.strings.count("two_temperature_file") > 0;
    const std::string tt_file = tt ? dir + s.strings.at("two_temperature_file") : std::string();
    const Status st = load_chemistry_ex(mech, therm, tran, nasa, tt_file, *tables, err);
    if (st != Status::Ok) {
      std::fprintf(stderr, "corefl_setup_load: %s\n", err.c_str());
      return static_cast<int>(st);
    }
    run->n_spec = tables->n_spec;
    run->two_temperature = tt ? 1 : 0;
    run->vt_relaxation = tt && get_or(s.bools, std::string("enable_vt_relaxation"), true) ? 1 : 0;
  }
  run->n_scalar = run->n_spec + run->two_temperature;
  run->n_scalar_transported = run->n_scalar;
  run->n_var = 5 + run->n_scalar;

  // reference state from the named inflow struct
  const std::string ref = get_or(s.strings, std::string("reference_state"), std::string(""));
  auto it = s.structs.find(ref);
  if (it != s.structs.end()) {
    const auto &st = it->second;
    const double T_in = struct_real(st, "temperature", -1), p = struct_real(st, "pressure", -1);
    double rho = struct_real(st, "density", -1), vel = struct_real(st, "velocity", 0);
    const double mach = struct_real(st, "mach", -1);
    double mw = 28.9647, gamma = 1.4;
    std::vector<double> y;
    if (run->mixture) {
      // species mass fractions are struct keys named after the species
      y.assign(tables->n_spec, 0.0);
      for (int l = 0; l < tables->n_spec; ++l) y[l] = struct_real(st, species_name(*tables, dir, s, l), 0.0);
      double imw = 0;
      for (int l = 0; l < tables->n_spec; ++l) imw += y[l] / tables->mw[l];
      mw = imw > 0 ? 1 / imw : 28.9647;
    }
    double T = T_in;
    if (T < 0 && rho > 0 && p > 0) T = p * mw / (rho * kRu);
    if (rho < 0 && p > 0 && T > 0) rho = p * mw / (kRu * T);
    const double R = kRu / mw;
    if (run->mixture && T > 0) {
      double cp = 0;
      for (int l = 0; l < tables->n_spec; ++l) cp += y[l] * host_cp(*tables, l, T);
      gamma = cp / (cp - R);
    }
    const double c = T > 0 ? std::sqrt(gamma * R * T) : 0;
    if (mach >= 0) vel = mach * c;
    run->rho_ref = rho > 0 ? rho : 1.0;
    run->v_ref = std::fabs(vel) < 1 ? c : vel;
    run->p_ref = p > 0 ? p : (rho > 0 && T > 0 ? rho * R * T : 1.0);
  } else {
    run->rho_ref = 1.0;
    run->v_ref = 1.0;
    run->p_ref = 1.0;
  }
  return 0;
}
