// Host-side Chemkin readers: mechanism (ELEMENTS/SPECIES/REACTIONS), NASA-7 or
// NASA-9 therm.dat, the transport file and the two-temperature metadata, producing the flat tables the device code
// reads.  Behaviour follows the reference readers in src/ChemData.cpp; the
// derived device tables follow src/DParameter.cu:80-88.
#include "corefl_internal.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace corefl {
namespace {

constexpr double kRc = 1.98720425864083;      // cal/(mol K), src/Constants.h:21
constexpr double kBoltzmann = 1.380649e-23;   // J/K, src/Constants.h:31
constexpr double kPi = 3.141592653589793;

// IUPAC standard atomic weights (the values src/Element.cpp:13-30 tabulates).
double atomic_weight(const std::string &el) {
  static const std::map<std::string, double> w{{"H", 1.008},   {"HE", 4.002602}, {"C", 12.011},
                                               {"N", 14.007},  {"O", 15.999},    {"AR", 39.95},
                                               {"NE", 20.1797}, {"F", 18.998403163}};
  auto it = w.find(el);
  return it == w.end() ? -1.0 : it->second;
}

std::string upper(std::string s) {
  for (auto &c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

std::string strip_comment(const std::string &s) {
  auto p = s.find('!');
  return p == std::string::npos ? s : s.substr(0, p);
}

struct Mech {
  std::vector<std::string> elems;               // file order
  std::vector<std::string> species;             // file order = species index
  std::map<std::string, int> spec_index;
};

Status read_elements_species(const std::string &path, Mech &m, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open mechanism file " + path;
    return Status::Io;
  }
  std::vector<std::string> toks;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ss(strip_comment(line));
    std::string t;
    while (ss >> t) toks.push_back(t);
  }
  size_t i = 0;
  while (i < toks.size() && toks[i] != "ELEMENTS" && toks[i] != "ELEM") ++i;
  ++i;
  for (; i < toks.size(); ++i) {
    const std::string u = upper(toks[i]);
    if (u == "END") continue;
    if (u == "SPECIES" || u == "SPEC") {
      ++i;
      break;
    }
    m.elems.push_back(u);
  }
  for (; i < toks.size(); ++i) {
    const std::string u = upper(toks[i]);
    if (u == "END") continue;
    if (u == "REACTIONS" || u == "REAC" || u == "THERMO") break;
    m.spec_index[u] = static_cast<int>(m.species.size());
    m.species.push_back(u);
  }
  if (m.species.empty()) {
    err = "no species found in " + path;
    return Status::Parse;
  }
  if (static_cast<int>(m.species.size()) > COREFL_MAX_SPEC) {
    err = "too many species for COREFL_MAX_SPEC";
    return Status::Arg;
  }
  return Status::Ok;
}

double field(const std::string &s, size_t pos, size_t len) {
  if (pos >= s.size()) return 0.0;
  return std::stod(s.substr(pos, len));
}

// NASA-7 two-range reader (Combustion2Part), src/ChemData.cpp:545-862.
Status read_therm(const std::string &path, const Mech &m, corefl_chem_tables &t, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open thermo file " + path;
    return Status::Io;
  }
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(f, line)) lines.push_back(line);
  size_t pos = 0;
  while (pos < lines.size() && upper(lines[pos]).rfind("THERMO", 0) != 0) ++pos;
  ++pos;
  double tl = 300, tm = 1000, th = 5000;
  for (; pos < lines.size(); ++pos) {
    if (lines[pos].empty() || lines[pos][0] == '!') continue;
    std::istringstream ss(lines[pos]);
    ss >> tl >> tm >> th;
    ++pos;
    break;
  }
  const int ns = static_cast<int>(m.species.size());
  for (int s = 0; s < ns; ++s) {
    t.t_low[s] = tl;
    t.t_mid[s] = tm;
    t.t_high[s] = th;
  }
  std::vector<bool> have(ns, false);
  int n_read = 0;
  while (pos < lines.size()) {
    const std::string s = upper(lines[pos]);
    ++pos;
    if (s.empty() || s[0] == '!') continue;
    std::string key;
    std::istringstream(s) >> key;
    if (key == "END") continue;
    if (key == "REACTIONS" || key == "REAC") break;
    if (n_read >= ns) break;
    std::string name;
    std::istringstream(s.substr(0, 18)) >> name;
    auto it = m.spec_index.find(name);
    if (it == m.spec_index.end() || have[it->second]) {
      pos += 3;
      continue;
    }
    if (pos + 3 > lines.size()) {
      err = "truncated thermo entry for " + name;
      return Status::Parse;
    }
    const int sp = it->second;
    try {
      t.t_low[sp] = field(s, 45, 10);
      t.t_high[sp] = field(s, 55, 10);
      if (s.size() > 65) {
        std::string mid;
        std::istringstream(s.substr(65, 10)) >> mid;
        if (!mid.empty()) t.t_mid[sp] = std::stod(mid);
      }
      std::map<std::string, int> comp;
      for (int e = 0; e < 4; ++e) {
        if (s.size() < static_cast<size_t>(24 + e * 5)) break;
        std::string c = s.substr(24 + e * 5, 5);
        c.erase(0, c.find_first_not_of(' '));
        if (c.empty() || c[0] == '0') break;
        std::istringstream cs(c);
        std::string el;
        int n = 0;
        cs >> el >> n;
        if (std::find(m.elems.begin(), m.elems.end(), el) == m.elems.end()) continue;
        comp[el] = n;
      }
      // molecular weight: sum in alphabetical element order (std::map), src/ChemData.cpp:706-711
      std::vector<std::string> sorted = m.elems;
      std::sort(sorted.begin(), sorted.end());
      double w = 0;
      for (const auto &el : sorted) {
        const double aw = atomic_weight(el);
        if (aw < 0) {
          err = "unknown element " + el;
          return Status::Parse;
        }
        auto c = comp.find(el);
        w += aw * (c == comp.end() ? 0 : c->second);
      }
      t.mw[sp] = w;
      const std::string &l2 = lines[pos], &l3 = lines[pos + 1], &l4 = lines[pos + 2];
      for (int k = 0; k < 5; ++k) t.high_coeff[sp][k] = field(l2, 15 * k, 15);
      t.high_coeff[sp][5] = field(l3, 0, 15);
      t.high_coeff[sp][6] = field(l3, 15, 15);
      t.low_coeff[sp][0] = field(l3, 30, 15);
      t.low_coeff[sp][1] = field(l3, 45, 15);
      t.low_coeff[sp][2] = field(l3, 60, 15);
      for (int k = 0; k < 4; ++k) t.low_coeff[sp][3 + k] = field(l4, 15 * k, 15);
    } catch (const std::exception &) {
      err = "malformed thermo entry for " + name;
      return Status::Parse;
    }
    pos += 3;
    have[sp] = true;
    ++n_read;
  }
  if (n_read < ns) {
    err = "thermo data missing for some species";
    return Status::Parse;
  }
  return Status::Ok;
}

// NASA-9 multi-range reader, src/ChemData.cpp:865-1034: per species a line
// with the range count (columns 0-1) and the molecular weight (columns 52-64),
// then per range the bounds (columns 0-10, 11-21) and two lines of 16-column
// coefficients a1..a5 / a6 a7 _ b1 b2 (Fortran 'D' exponents allowed).
Status read_therm9(const std::string &path, const Mech &m, corefl_chem_tables &t, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open thermo file " + path;
    return Status::Io;
  }
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(f, line)) {
    std::replace(line.begin(), line.end(), 'D', 'E');
    std::replace(line.begin(), line.end(), 'd', 'e');
    lines.push_back(line);
  }
  size_t pos = 0;
  while (pos < lines.size() && upper(lines[pos]).find("THERMO") == std::string::npos) ++pos;
  ++pos;
  // header line of default bounds (T_low T_mid T_high T_hyper): read and unused
  for (; pos < lines.size(); ++pos) {
    if (lines[pos].empty() || lines[pos][0] == '!') continue;
    ++pos;
    break;
  }
  const int ns = static_cast<int>(m.species.size());
  std::vector<bool> have(ns, false);
  int n_read = 0;
  auto n_parts = [&](size_t at) {
    return at < lines.size() && lines[at].size() >= 2 ? std::atoi(lines[at].substr(0, 2).c_str()) : -1;
  };
  while (pos < lines.size()) {
    const std::string s = upper(lines[pos]);
    ++pos;
    if (s.empty() || s[0] == '!') continue;
    std::string key;
    std::istringstream(s) >> key;
    if (key == "END") continue;
    if (key == "REACTIONS" || key == "REAC" || key == "TRANSPORT" || key == "TRAN" || key == "TRANS") break;
    if (n_read >= ns) continue;
    std::string name;
    std::istringstream(s.substr(0, 16)) >> name;
    const int np = n_parts(pos);
    if (np < 1 || np > COREFL_MAX_RANGE) {
      err = "malformed NASA-9 entry for " + name;
      return Status::Parse;
    }
    auto it = m.spec_index.find(name);
    if (it == m.spec_index.end() || have[it->second]) {
      pos += 1 + 3 * static_cast<size_t>(np);
      continue;
    }
    if (pos + 1 + 3 * static_cast<size_t>(np) > lines.size()) {
      err = "truncated NASA-9 entry for " + name;
      return Status::Parse;
    }
    const int sp = it->second;
    try {
      t.mw[sp] = field(lines[pos], 52, 13);
      t.n_range[sp] = np;
      for (int p = 0; p < np; ++p) {
        const std::string &lb = lines[pos + 1 + 3 * p], &l1 = lines[pos + 2 + 3 * p], &l2 = lines[pos + 3 + 3 * p];
        t.t_cut[sp][p] = field(lb, 0, 11);
        t.t_cut[sp][p + 1] = field(lb, 11, 11);
        for (int k = 0; k < 5; ++k) t.poly9[sp][p][k] = field(l1, 16 * k, 16);
        t.poly9[sp][p][5] = field(l2, 0, 16);
        t.poly9[sp][p][6] = field(l2, 16, 16);
        t.poly9[sp][p][7] = field(l2, 48, 16);
        t.poly9[sp][p][8] = field(l2, 64, 16);
      }
    } catch (const std::exception &) {
      err = "malformed NASA-9 entry for " + name;
      return Status::Parse;
    }
    pos += 1 + 3 * static_cast<size_t>(np);
    have[sp] = true;
    ++n_read;
  }
  if (n_read < ns) {
    err = "NASA-9 thermo data missing for some species";
    return Status::Parse;
  }
  return Status::Ok;
}

// Two-temperature metadata, src/ChemData.cpp:130-232: SPECIES <name> ... END
// blocks with MODE, THETA_V, PARK_SIGMA (or SIGMA / OMEGA), ELEC_THETA and
// ELEC_G lists and LT <partner> a b lines; '#' and '!' start comments;
// species absent from the mechanism are skipped.
Status read_two_temperature(const std::string &path, const Mech &m, corefl_chem_tables &t, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open two-temperature file " + path;
    return Status::Io;
  }
  const int ns = static_cast<int>(m.species.size());
  std::vector<std::vector<double>> theta(ns), g(ns);
  for (int i = 0; i < ns; ++i) {
    t.theta_v[i] = 0.0;
    t.park_sigma[i] = 3.0e-21;
    for (int j = 0; j < ns; ++j) t.lt_a[i][j] = t.lt_b[i][j] = 0.0;
  }
  std::string line;
  int cur = -1;
  while (std::getline(f, line)) {
    const auto c = line.find_first_of("#!");
    if (c != std::string::npos) line.erase(c);
    std::istringstream ss(line);
    std::string key;
    if (!(ss >> key)) continue;
    key = upper(key);
    if (key == "SPECIES") {
      std::string name;
      ss >> name;
      auto it = m.spec_index.find(upper(name));
      cur = it == m.spec_index.end() ? -2 : it->second;
      continue;
    }
    if (key == "END") {
      cur = -1;
      continue;
    }
    if (cur < 0) continue;
    if (key == "THETA_V") {
      ss >> t.theta_v[cur];
    } else if (key == "PARK_SIGMA" || key == "SIGMA" || key == "OMEGA") {
      ss >> t.park_sigma[cur];
    } else if (key == "ELEC_THETA") {
      double v;
      while (ss >> v) theta[cur].push_back(v);
    } else if (key == "ELEC_G") {
      double v;
      while (ss >> v) g[cur].push_back(v);
    } else if (key == "LT") {
      std::string partner;
      double a = 0, b = 0;
      ss >> partner >> a >> b;
      auto it = m.spec_index.find(upper(partner));
      if (it == m.spec_index.end()) continue;
      t.lt_a[cur][it->second] = a;
      t.lt_b[cur][it->second] = b;
    }
  }
  for (int i = 0; i < ns; ++i) {
    if (g[i].empty() && !theta[i].empty()) g[i].assign(theta[i].size(), 1.0);
    if (theta[i].size() != g[i].size()) {
      err = "electronic level count mismatch for species " + m.species[i];
      return Status::Parse;
    }
    if (theta[i].size() > COREFL_MAX_ELEC) {
      err = "more electronic levels than COREFL_MAX_ELEC for species " + m.species[i];
      return Status::Arg;
    }
    t.n_elec[i] = static_cast<int>(theta[i].size());
    for (int j = 0; j < t.n_elec[i]; ++j) {
      t.elec_theta[i][j] = theta[i][j];
      t.elec_g[i][j] = g[i][j];
    }
  }
  t.two_temperature = 1;
  return Status::Ok;
}

double reduced_dipole(double mu, double eps_kb, double sigma) {
  if (mu == 0) return 0;
  const double debye_to_cm_joule = 1e-18 * std::sqrt(1e-7);
  const double cm_to_angstrom = std::pow(1e+8, 1.5);
  const double convert = debye_to_cm_joule * cm_to_angstrom / std::sqrt(kBoltzmann);
  return convert * mu / std::sqrt(eps_kb * sigma * sigma * sigma);
}

// Transport reader and derived tables, src/ChemData.cpp:1037-1162.
Status read_transport(const std::string &path, const Mech &m, corefl_chem_tables &t, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open transport file " + path;
    return Status::Io;
  }
  const int ns = static_cast<int>(m.species.size());
  std::vector<double> eps(ns, 0), sig(ns, 0), mu(ns, 0), alp(ns, 0), zr(ns, 0);
  std::vector<bool> have(ns, false);
  int n_read = 0;
  std::string line;
  while (std::getline(f, line)) {
    const std::string s = upper(line);
    if (s.find_first_not_of(" \t\r") == std::string::npos || s[0] == '!') continue;
    std::istringstream ss(s);
    std::string key;
    ss >> key;
    if (key.rfind("END", 0) == 0 || key.rfind("REAC", 0) == 0) break;
    auto it = m.spec_index.find(key);
    if (it == m.spec_index.end()) continue;
    if (n_read >= ns) break;
    const int sp = it->second;
    if (have[sp]) continue;
    int geom = 0;
    if (!(ss >> geom >> eps[sp] >> sig[sp] >> mu[sp] >> alp[sp] >> zr[sp])) {
      err = "malformed transport entry for " + key;
      return Status::Parse;
    }
    t.geometry[sp] = geom;
    have[sp] = true;
    ++n_read;
  }
  if (n_read < ns) {
    err = "transport data missing for some species";
    return Status::Parse;
  }
  for (int s = 0; s < ns; ++s) {
    t.LJ_potent_inv[s] = 1.0 / eps[s];
    t.vis_coeff[s] = 2.6693e-6 * std::sqrt(t.mw[s]) / (sig[s] * sig[s]);
    const double tri = eps[s] / 298;
    // Parker F(298) as the reference writes it (src/ChemData.cpp:1083-1085)
    const double f298 = 1 + 0.5 * kPi * kPi * kPi * std::sqrt(tri) + (2 + 0.25 * kPi * kPi) * tri +
                        std::sqrt(kPi * kPi * kPi) * std::sqrt(tri * tri * tri);
    t.ZRotF298[s] = zr[s] * f298;
  }
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      t.WjDivWi_to_One4th[i][j] = std::pow(t.mw[j] / t.mw[i], 0.25);
      t.sqrt_WiDivWjPl1Mul8[i][j] = 1.0 / std::sqrt(8 * (1 + t.mw[i] / t.mw[j]));
    }
  for (int j = 0; j < ns; ++j) {
    const double wjj = t.mw[j] * 0.5;
    t.binary_diffusivity_coeff[j][j] = 0.0188324 / (std::sqrt(wjj) * sig[j] * sig[j]);
    t.kb_over_eps_jk[j][j] = 1.0 / eps[j];
    for (int k = j + 1; k < ns; ++k) {
      const double wjk = t.mw[j] * t.mw[k] / (t.mw[j] + t.mw[k]);
      double xi = 1.0;
      if ((mu[j] != 0) != (mu[k] != 0)) {
        const int p = mu[j] != 0 ? j : k, n = mu[j] != 0 ? k : j;
        const double alpha_n_star = alp[n] / std::pow(sig[n], 3);
        const double mu_p_star = reduced_dipole(mu[p], eps[p], sig[p]);
        xi = 1.0 + 0.25 * alpha_n_star * mu_p_star * mu_p_star * std::sqrt(eps[p] / eps[n]);
      }
      const double sjk = 0.5 * (sig[j] + sig[k]) * std::pow(xi, -1.0 / 6.0);
      const double bd = 0.0188324 / (std::sqrt(wjk) * sjk * sjk);
      t.binary_diffusivity_coeff[j][k] = t.binary_diffusivity_coeff[k][j] = bd;
      const double e = xi * xi * std::sqrt(eps[j] * eps[k]);
      t.kb_over_eps_jk[j][k] = t.kb_over_eps_jk[k][j] = 1.0 / e;
    }
  }
  return Status::Ok;
}

struct Rxn {
  int label = 1, rev = 0, order = 0;
  std::vector<int> sf, sb;
  double A = 0, b = 0, Ea = 0, A2 = 0, b2 = 0, Ea2 = 0;
  std::vector<double> tb;
  double ta = 0, t1 = 0, t2 = 0, t3 = 0;
  bool dup = false;
};

// '+'-separated species terms with an optional one-digit coefficient
// (src/ChemData.cpp:1380-1437).
Status add_side(const std::string &side, const Mech &m, std::vector<int> &row, int sign, int &order,
                std::string &err) {
  std::vector<std::string> terms;
  std::string cur;
  for (size_t i = 0; i < side.size(); ++i) {
    const char c = side[i];
    if (c == '+' && !cur.empty() && !(i + 1 < side.size() && side[i + 1] == '+')) {
      terms.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (!cur.empty()) terms.push_back(cur);
  for (auto term : terms) {
    int st = 1;
    if (!term.empty() && std::isdigit(static_cast<unsigned char>(term[0]))) {
      st = term[0] - '0';
      term.erase(0, 1);
    }
    auto it = m.spec_index.find(term);
    if (it == m.spec_index.end()) {
      err = "unknown species '" + term + "' in reaction";
      return Status::Parse;
    }
    row[it->second] += st;
    order += sign * st;
  }
  return Status::Ok;
}

void erase_all(std::string &s, const std::string &what, int max_count) {
  for (int n = 0; n < max_count; ++n) {
    auto p = s.find(what);
    if (p == std::string::npos) return;
    s.erase(p, what.size());
  }
}

// Reaction reader, src/ChemData.cpp:1165-1493.
Status read_reactions(const std::string &path, const Mech &m, corefl_chem_tables &t, std::string &err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open mechanism file " + path;
    return Status::Io;
  }
  const int ns = static_cast<int>(m.species.size());
  std::string line;
  bool in_reac = false;
  std::vector<Rxn> all;
  while (std::getline(f, line)) {
    const std::string s = upper(line);
    if (!in_reac) {
      std::string key;
      std::istringstream(s) >> key;
      if (key == "REACTIONS" || key == "REAC") in_reac = true;
      continue;
    }
    std::string trimmed = s;
    trimmed.erase(0, trimmed.find_first_not_of(" \t\r"));
    if (trimmed.rfind("END", 0) == 0 && trimmed.find('=') == std::string::npos) break;
    if (s.empty() || s[0] == '!') continue;
    std::string body = strip_comment(s);
    if (body.find_first_not_of(" \t\r") == std::string::npos) continue;
    if (body.find('=') != std::string::npos) {
      std::istringstream ss(body);
      std::vector<std::string> tok;
      std::string x;
      while (ss >> x) tok.push_back(x);
      if (tok.size() < 4) {
        err = "incomplete reaction line: " + line;
        return Status::Parse;
      }
      Rxn r;
      r.sf.assign(ns, 0);
      r.sb.assign(ns, 0);
      r.tb.assign(ns, 1.0);
      std::string eq;
      for (size_t k = 0; k + 3 < tok.size(); ++k) eq += tok[k];
      try {
        r.A = std::stod(tok[tok.size() - 3]);
        r.b = std::stod(tok[tok.size() - 2]);
        r.Ea = std::stod(tok[tok.size() - 1]) / kRc;
      } catch (const std::exception &) {
        err = "bad Arrhenius parameters: " + line;
        return Status::Parse;
      }
      if (eq.find("+M") != std::string::npos) {
        r.label = 4;
        if (eq.find("(+") != std::string::npos) {
          r.label = 5;
          erase_all(eq, "(+M)", 2);
        } else {
          erase_all(eq, "+M", 2);
        }
      }
      std::string lhs, rhs;
      size_t p;
      if ((p = eq.find("<=>")) != std::string::npos) {
        lhs = eq.substr(0, p);
        rhs = eq.substr(p + 3);
      } else if ((p = eq.find("=>")) != std::string::npos) {
        lhs = eq.substr(0, p);
        rhs = eq.substr(p + 2);
        r.label = 0;
      } else {
        p = eq.find('=');
        lhs = eq.substr(0, p);
        rhs = eq.substr(p + 1);
      }
      Status st = add_side(lhs, m, r.sf, -1, r.order, err);
      if (st != Status::Ok) return st;
      st = add_side(rhs, m, r.sb, +1, r.order, err);
      if (st != Status::Ok) return st;
      all.push_back(std::move(r));
      continue;
    }
    if (all.empty()) continue;
    Rxn &r = all.back();
    std::string aux = body;
    std::replace(aux.begin(), aux.end(), '/', ' ');
    std::istringstream ss(aux);
    std::vector<std::string> tok;
    std::string x;
    while (ss >> x) tok.push_back(x);
    if (tok.empty()) continue;
    try {
      if (tok[0] == "DUPLICATE" || tok[0] == "DUP") {
        r.label = 3;
        r.dup = true;
      } else if (tok[0] == "REV" && tok.size() >= 4) {
        r.A2 = std::stod(tok[1]);
        r.b2 = std::stod(tok[2]);
        r.Ea2 = std::stod(tok[3]) / kRc;
        r.rev = 1;
      } else if (tok[0] == "LOW" && tok.size() >= 4) {
        r.A2 = std::stod(tok[1]);
        r.b2 = std::stod(tok[2]);
        r.Ea2 = std::stod(tok[3]) / kRc;
      } else if (tok[0] == "TROE" && tok.size() >= 4) {
        r.label = 6;
        r.ta = std::stod(tok[1]);
        r.t3 = std::stod(tok[2]);
        r.t1 = std::stod(tok[3]);
        if (tok.size() > 4) {
          r.t2 = std::stod(tok[4]);
          r.label = 7;
        }
      } else if (tok[0] == "PLOG") {
        continue;
      } else {
        for (size_t k = 0; k + 1 < tok.size(); k += 2) {
          auto it = m.spec_index.find(tok[k]);
          if (it == m.spec_index.end()) {
            err = "unknown third-body species " + tok[k];
            return Status::Parse;
          }
          r.tb[it->second] = std::stod(tok[k + 1]);
        }
      }
    } catch (const std::exception &) {
      err = "malformed auxiliary line: " + line;
      return Status::Parse;
    }
  }
  // DUPLICATE pairs: the second's Arrhenius set becomes A2/b2/Ea2 of the first
  // and the second is dropped (src/ChemData.cpp:1206-1243).
  std::vector<Rxn> out;
  std::vector<size_t> pending;
  for (auto &r : all) {
    if (r.dup) {
      bool merged = false;
      for (size_t k = 0; k < pending.size(); ++k) {
        Rxn &prev = out[pending[k]];
        if (prev.sf == r.sf && prev.sb == r.sb) {
          prev.A2 = r.A;
          prev.b2 = r.b;
          prev.Ea2 = r.Ea;
          pending.erase(pending.begin() + static_cast<long>(k));
          merged = true;
          break;
        }
      }
      if (merged) continue;
      pending.push_back(out.size());
    }
    out.push_back(r);
  }
  if (static_cast<int>(out.size()) > COREFL_MAX_REAC) {
    err = "too many reactions for COREFL_MAX_REAC";
    return Status::Arg;
  }
  t.n_reac = static_cast<int>(out.size());
  for (int i = 0; i < t.n_reac; ++i) {
    const Rxn &r = out[i];
    t.reac_type[i] = r.label;
    t.rev_type[i] = r.rev;
    t.order[i] = r.order;
    t.A[i] = r.A;
    t.b[i] = r.b;
    t.Ea[i] = r.Ea;
    t.A2[i] = r.A2;
    t.b2[i] = r.b2;
    t.Ea2[i] = r.Ea2;
    t.troe_alpha[i] = r.ta;
    t.troe_t1[i] = r.t1;
    t.troe_t2[i] = r.t2;
    t.troe_t3[i] = r.t3;
    for (int s = 0; s < ns; ++s) {
      t.stoi_f[i][s] = r.sf[s];
      t.stoi_b[i][s] = r.sb[s];
      t.third_body[i][s] = r.tb[s];
    }
  }
  return Status::Ok;
}

}  // namespace

Status load_chemistry(const std::string &mech, const std::string &therm, const std::string &tran,
                      corefl_chem_tables &t, std::string &err) {
  return load_chemistry_ex(mech, therm, tran, 7, "", t, err);
}

Status load_chemistry_ex(const std::string &mech, const std::string &therm, const std::string &tran, int nasa,
                         const std::string &two_temperature, corefl_chem_tables &t, std::string &err) {
  std::memset(&t, 0, sizeof(t));
  if (nasa != 7 && nasa != 9) {
    err = "therm_nasa_7_9 must be 7 or 9";
    return Status::Arg;
  }
  Mech m;
  Status st = read_elements_species(mech, m, err);
  if (st != Status::Ok) return st;
  t.n_spec = static_cast<int>(m.species.size());
  for (int s = 0; s < t.n_spec; ++s) {
    if (m.species[s].size() >= COREFL_SPEC_NAME_LEN) {
      err = "species name longer than " + std::to_string(COREFL_SPEC_NAME_LEN - 1) + " characters: " + m.species[s];
      return Status::Arg;
    }
    std::snprintf(t.species_name[s], COREFL_SPEC_NAME_LEN, "%s", m.species[s].c_str());
  }
  t.nasa7or9 = nasa;
  st = nasa == 9 ? read_therm9(therm, m, t, err) : read_therm(therm, m, t, err);
  if (st != Status::Ok) return st;
  for (int s = 0; s < t.n_spec; ++s) {
    t.imw[s] = 1.0 / t.mw[s];
    t.gas_const[s] = kRu / t.mw[s];
  }
  if ((st = read_transport(tran, m, t, err)) != Status::Ok) return st;
  if (!two_temperature.empty() && (st = read_two_temperature(two_temperature, m, t, err)) != Status::Ok) return st;
  if ((st = read_reactions(mech, m, t, err)) != Status::Ok) return st;
  // reaction temperatures: T for every reaction (a = 1, b = 0), src/ChemData.cpp:1294-1297; with
  // two-temperature data, the 17-reaction N2/O2/NO/N/O mechanism takes sqrt(T Tve) for its first
  // 15 (dissociation) reactions, src/ChemData.cpp:1248-1262
  for (int r = 0; r < t.n_reac; ++r) {
    t.tcf_a[r] = t.tcb_a[r] = 1.0;
    t.tcf_b[r] = t.tcb_b[r] = 0.0;
  }
  if (t.two_temperature && t.n_spec == 5 && t.n_reac == 17) {
    bool air5 = true;
    for (const char *nm : {"N2", "O2", "NO", "N", "O"}) air5 = air5 && m.spec_index.count(nm);
    if (air5) {
      t.trx_model = 1;
      for (int r = 0; r < 15; ++r) t.tcf_a[r] = t.tcf_b[r] = 0.5;
    }
  }
  return Status::Ok;
}

}  // namespace corefl
