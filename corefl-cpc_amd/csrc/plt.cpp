// Tecplot binary (V112) solution files of a context's blocks: the reference's
// output/flowfield.plt writer (FieldIO<>::write_header, compute_offset_header
// and write_common_data_section, src/FieldIO.h:96-420; print_field 709-800
// writes the same data section again for later outputs) and its restart reader
// read_flowfield (src/Initialize.cu:167-420), with output/message/step.txt
// holding the step count.  Host code over the context's field download and
// upload; every value in the file is FP64.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "corefl_internal.h"

namespace {

using corefl::Status;

// Little-endian stream writer with the file's primitive encodings: a string
// is one int32 per character and a terminating 0 (gxl::write_str).
class Out {
 public:
  explicit Out(const char *path) : f_(std::fopen(path, "wb")) {}
  ~Out() {
    if (f_) std::fclose(f_);
  }
  bool open() const { return f_ != nullptr; }
  bool ok() const { return f_ && !bad_; }
  bool close() {
    const bool good = ok() && std::fclose(f_) == 0;
    f_ = nullptr;
    return good;
  }
  void bytes(const void *p, size_t n) {
    if (f_ && std::fwrite(p, 1, n, f_) != n) bad_ = true;
  }
  void i32(int32_t v) { bytes(&v, 4); }
  void f32(float v) { bytes(&v, 4); }
  void f64(double v) { bytes(&v, 8); }
  void str(const std::string &s) {
    for (unsigned char ch : s) i32(static_cast<int32_t>(ch));
    i32(0);
  }

 private:
  std::FILE *f_;
  bool bad_ = false;
};

class In {
 public:
  explicit In(const char *path) : f_(std::fopen(path, "rb")) {}
  ~In() {
    if (f_) std::fclose(f_);
  }
  bool open() const { return f_ != nullptr; }
  bool ok() const { return f_ && !bad_; }
  void bytes(void *p, size_t n) {
    if (!f_ || std::fread(p, 1, n, f_) != n) bad_ = true;
  }
  int32_t i32() {
    int32_t v = 0;
    bytes(&v, 4);
    return v;
  }
  float f32() {
    float v = 0;
    bytes(&v, 4);
    return v;
  }
  double f64() {
    double v = 0;
    bytes(&v, 8);
    return v;
  }
  std::string str() {
    std::string s;
    for (int n = 0; n < 4096; ++n) {
      const int32_t v = i32();
      if (!ok() || v == 0) return s;
      s += static_cast<char>(v);
    }
    bad_ = true;
    return s;
  }
  void skip(long long n) {
    if (!f_ || fseeko(f_, static_cast<off_t>(n), SEEK_CUR) != 0) bad_ = true;
  }

 private:
  std::FILE *f_;
  bool bad_ = false;
};

std::string upper(std::string s) {
  for (char &ch : s) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
  return s;
}

// The variables in the reference's order (Parameter::get_variable_names,
// src/Parameter.cpp:835-945, for a laminar run without passive scalars):
// coordinates, bv, the scalars (species, then Eve with two temperatures), then
// mach (and Tve).
std::vector<std::string> variable_names(const corefl::Run &r, const corefl_chem_tables &t) {
  std::vector<std::string> v{"x", "y", "z", "density", "u", "v", "w", "pressure", "temperature"};
  for (int l = 0; l < r.ns; ++l) {
    const char *nm = t.species_name[l];
    v.emplace_back(nm[0] ? std::string(nm, strnlen(nm, COREFL_SPEC_NAME_LEN)) : "Y" + std::to_string(l));
  }
  if (r.tt) v.emplace_back("Eve");
  v.emplace_back("mach");
  if (r.tt) v.emplace_back("Tve");
  return v;
}

// One padded plane's (file) extent, i fastest: the block's nodes and ngg_out of its ghost layers.
void extract(const double *plane, const corefl::BlockExtent &b, int g, std::vector<double> &out) {
  const long px = b.mx + 2L * b.ngg, pxy = px * (b.my + 2L * b.ngg);
  const int nx = b.mx + 2 * g, ny = b.my + 2 * g, nz = b.mz + 2 * g, o = b.ngg - g;
  out.resize(static_cast<size_t>(nx) * ny * nz);
  size_t n = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const double *row = plane + (k + o) * pxy + (j + o) * px + o;
      std::memcpy(&out[n], row, sizeof(double) * nx);
      n += nx;
    }
}

void insert(const std::vector<double> &in, const corefl::BlockExtent &b, int g, double *plane) {
  const long px = b.mx + 2L * b.ngg, pxy = px * (b.my + 2L * b.ngg);
  const int nx = b.mx + 2 * g, ny = b.my + 2 * g, nz = b.mz + 2 * g, o = b.ngg - g;
  size_t n = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      std::memcpy(plane + (k + o) * pxy + (j + o) * px + o, &in[n], sizeof(double) * nx);
      n += nx;
    }
}

}  // namespace

extern "C" int corefl_plt_write(corefl_ctx *ctx, const char *path, const double *const *coords, int ngg_out,
                                double solution_time, const char *step_path, int step) {
  using namespace corefl;
  if (!ctx) return static_cast<int>(Status::Arg);
  const CtxView cv = ctx_view(ctx);
  if (!path) return ctx_fail(ctx, Status::Arg, "corefl_plt_write: no path");
  for (const BlockExtent &b : cv.blocks)
    if (ngg_out < 0 || ngg_out > b.ngg) return ctx_fail(ctx, Status::Arg, "corefl_plt_write: ngg_out outside 0..ngg");
  const Run &r = *cv.run;
  const std::vector<std::string> names = variable_names(r, *cv.tables);
  const int n_var = static_cast<int>(names.size());
  const int nb = static_cast<int>(cv.blocks.size());

  // one block's planes in the file's variable order, over the file extent (the file is written a block
  // at a time, so host memory holds one block)
  std::vector<std::vector<double>> d;
  std::vector<double> host, part;
  auto gather = [&](int b) -> int {
    const BlockExtent &e = cv.blocks[b];
    const long plane = (e.mx + 2L * e.ngg) * (e.my + 2L * e.ngg) * (e.mz + 2L * e.ngg);
    d.clear();
    for (int a = 0; a < 3; ++a) {
      if (coords && coords[b]) {
        extract(coords[b] + a * plane, e, ngg_out, part);
      } else {
        part.assign(static_cast<size_t>(e.mx + 2 * ngg_out) * (e.my + 2 * ngg_out) * (e.mz + 2 * ngg_out), 0.0);
      }
      d.push_back(part);
    }
    auto planes = [&](int field, int count) -> int {
      host.resize(static_cast<size_t>(plane) * count);
      const int st = corefl_field_download(ctx, b, field, host.data());
      if (st != 0) return st;
      for (int l = 0; l < count; ++l) {
        extract(host.data() + l * plane, e, ngg_out, part);
        d.push_back(part);
      }
      return 0;
    };
    int st = planes(COREFL_F_BV, 6);
    if (st == 0 && r.n_scalar > 0) st = planes(COREFL_F_SV, r.n_scalar);
    if (st == 0) st = planes(COREFL_F_MACH, 1);
    if (st == 0 && r.tt) st = planes(COREFL_F_TVE, 1);
    if (st != 0) return st;
    if (static_cast<int>(d.size()) != n_var)
      return ctx_fail(ctx, Status::Arg, "corefl_plt_write: scalar count does not match the variable names");
    return 0;
  };

  Out o(path);
  if (!o.open()) return ctx_fail(ctx, Status::Arg, std::string("corefl_plt_write: cannot create ") + path);
  // I. header: magic number, byte order, file type (0 full), title, variable names
  o.bytes("#!TDV112", 8);
  o.i32(1);
  o.i32(0);
  o.str("Solution file");
  o.i32(n_var);
  for (const auto &n : names) o.str(n);
  for (int b = 0; b < nb; ++b) {
    const BlockExtent &e = cv.blocks[b];
    o.f32(299.0f);  // zone marker
    o.str("zone " + std::to_string(b));
    o.i32(-1);  // parent zone
    o.i32(-2);  // strand id: pending
    o.f64(solution_time);
    o.i32(-1);  // zone colour
    o.i32(0);   // ordered
    o.i32(0);   // data at the nodes
    o.i32(0);   // no raw face neighbours
    o.i32(0);   // no miscellaneous face connections
    o.i32(e.mx + 2 * ngg_out);
    o.i32(e.my + 2 * ngg_out);
    o.i32(e.mz + 2 * ngg_out);
    o.i32(0);  // no auxiliary data
  }
  o.f32(357.0f);  // end of header
  // II. data section: per zone the formats (2 = double), no passive or shared variables, the min / max of
  // every variable, then the variables one after another
  for (int b = 0; b < nb; ++b) {
    if (const int st = gather(b)) return st;
    o.f32(299.0f);
    for (int l = 0; l < n_var; ++l) o.i32(2);
    o.i32(0);
    o.i32(0);
    o.i32(-1);
    for (const auto &v : d) {
      double lo = v.empty() ? 0.0 : v[0], hi = lo;
      for (double x : v) {
        lo = std::min(lo, x);
        hi = std::max(hi, x);
      }
      o.f64(lo);
      o.f64(hi);
    }
    for (const auto &v : d) o.bytes(v.data(), v.size() * sizeof(double));
  }
  if (!o.close()) return ctx_fail(ctx, Status::Arg, std::string("corefl_plt_write: writing ") + path + " failed");
  if (step_path) {
    std::FILE *f = std::fopen(step_path, "w");
    if (!f) return ctx_fail(ctx, Status::Arg, std::string("corefl_plt_write: cannot create ") + step_path);
    std::fprintf(f, "%d", step);
    if (std::fclose(f) != 0) return ctx_fail(ctx, Status::Arg, std::string("corefl_plt_write: writing ") + step_path);
  }
  return 0;
}

extern "C" int corefl_plt_read(corefl_ctx *ctx, const char *path, double *solution_time) {
  using namespace corefl;
  if (!ctx) return static_cast<int>(Status::Arg);
  const CtxView cv = ctx_view(ctx);
  if (!path) return ctx_fail(ctx, Status::Arg, "corefl_plt_read: no path");
  const Run &r = *cv.run;
  const int nb = static_cast<int>(cv.blocks.size());
  In in(path);
  if (!in.open()) return ctx_fail(ctx, Status::Arg, std::string("corefl_plt_read: cannot open ") + path);
  auto bad = [&](const std::string &m) { return ctx_fail(ctx, Status::Arg, "corefl_plt_read: " + m); };
  char magic[8];
  in.bytes(magic, 8);
  if (!in.ok() || std::memcmp(magic, "#!TDV", 5) != 0) return bad("not a Tecplot binary file");
  if (in.i32() != 1) return bad("byte order other than 1");
  in.i32();  // file type
  in.str();  // title
  const int n_var = in.i32();
  if (!in.ok() || n_var < 3 || n_var > 4096) return bad("bad variable count");
  std::vector<std::string> names(n_var);
  for (auto &n : names) n = in.str();
  // labels as identify_variable_labels (src/Initialize.cu:424-520): bv 0..5, scalar l -> 6 + l, Tve -> -10,
  // anything else (coordinates, mach, ...) skipped
  std::vector<int> label(n_var, -1);
  std::vector<bool> have_scalar(r.n_scalar, false);
  for (int l = 0; l < n_var; ++l) {
    const std::string u = upper(names[l]);
    if (u == "RHO" || u == "DENSITY") label[l] = 0;
    else if (u == "U") label[l] = 1;
    else if (u == "V") label[l] = 2;
    else if (u == "W") label[l] = 3;
    else if (u == "P" || u == "PRESSURE") label[l] = 4;
    else if (u == "T" || u == "TEMPERATURE") label[l] = 5;
    else if (r.tt && u == "EVE") label[l] = 6 + r.ns;
    else if (r.tt && u == "TVE") label[l] = -10;
    else
      for (int s = 0; s < r.ns; ++s) {
        const char *nm = cv.tables->species_name[s];
        if (nm[0] && upper(std::string(nm, strnlen(nm, COREFL_SPEC_NAME_LEN))) == u) label[l] = 6 + s;
      }
    if (label[l] >= 6) have_scalar[label[l] - 6] = true;
  }
  for (int b = 0; b < 6; ++b)
    if (std::find(label.begin(), label.end(), b) == label.end()) return bad("variable of bv " + std::to_string(b) + " missing");
  for (int s = 0; s < r.n_scalar; ++s)
    if (!have_scalar[s]) return bad("scalar " + std::to_string(s) + " missing (species are matched by name)");
  // zones
  std::vector<std::array<int, 3>> dims;
  double time = 0;
  float marker = in.f32();
  while (in.ok() && std::abs(marker - 299.0f) < 1e-3f) {
    in.str();    // zone name
    in.skip(8);  // parent zone, strand id
    time = in.f64();
    in.skip(20);  // colour, zone type, data location, face neighbours, face connections
    std::array<int, 3> d{in.i32(), in.i32(), in.i32()};
    dims.push_back(d);
    if (in.i32() != 0) return bad("auxiliary zone data is not supported");
    marker = in.f32();
  }
  if (!in.ok() || std::abs(marker - 357.0f) > 1e-3f) return bad("malformed header");
  const int nz = static_cast<int>(dims.size());
  // zones to blocks: in order when every zone has its block's nodes, else a one-to-one match of node counts
  // (read_flowfield's same_mesh and same_block cases); anything else is an error
  std::vector<int> zone_of(nb, -1);
  auto same = [&](int b, int z) {
    const BlockExtent &e = cv.blocks[b];
    return dims[z][0] == e.mx && dims[z][1] == e.my && dims[z][2] == e.mz;
  };
  if (nz != nb) return bad(std::to_string(nz) + " zones for " + std::to_string(nb) + " blocks");
  bool in_order = true;
  for (int b = 0; b < nb; ++b) in_order = in_order && same(b, b);
  for (int b = 0; b < nb; ++b) {
    if (in_order) {
      zone_of[b] = b;
      continue;
    }
    int found = 0;
    for (int z = 0; z < nz; ++z)
      if (same(b, z)) {
        ++found;
        zone_of[b] = z;
      }
    if (found != 1) return bad("the zones do not match the blocks (interpolation is not supported)");
  }
  // data sections in file order; each zone's variables go to its block
  std::vector<int> block_of(nz, -1);
  for (int b = 0; b < nb; ++b) block_of[zone_of[b]] = b;
  std::vector<double> part, bvh, svh, tveh;
  for (int z = 0; z < nz; ++z) {
    const int b = block_of[z];
    const BlockExtent &e = cv.blocks[b];
    const long plane = (e.mx + 2L * e.ngg) * (e.my + 2L * e.ngg) * (e.mz + 2L * e.ngg);
    const long long n = static_cast<long long>(dims[z][0]) * dims[z][1] * dims[z][2];
    if (std::abs(in.f32() - 299.0f) > 1e-3f) return bad("malformed zone data");
    for (int l = 0; l < n_var; ++l)
      if (in.i32() != 2) return bad("variables other than double precision are not supported");
    if (in.i32() != 0 || in.i32() != 0) return bad("passive or shared variables are not supported");
    in.i32();                 // connectivity sharing
    in.skip(16LL * n_var);    // min / max pairs
    // the block's current padded planes: the file's extent is overwritten, the ghosts kept
    bvh.resize(static_cast<size_t>(plane) * 6);
    svh.resize(static_cast<size_t>(plane) * r.n_scalar);
    int st = corefl_field_download(ctx, b, COREFL_F_BV, bvh.data());
    if (st == 0 && r.n_scalar > 0) st = corefl_field_download(ctx, b, COREFL_F_SV, svh.data());
    if (st == 0 && r.tt) {
      tveh.resize(static_cast<size_t>(plane));
      st = corefl_field_download(ctx, b, COREFL_F_TVE, tveh.data());
    }
    if (st != 0) return st;
    for (int l = 0; l < n_var; ++l) {
      if (label[l] == -1) {
        in.skip(n * 8);
        continue;
      }
      part.resize(static_cast<size_t>(n));
      in.bytes(part.data(), static_cast<size_t>(n) * 8);
      if (!in.ok()) return bad("file ends inside the data");
      double *dst = label[l] == -10 ? tveh.data()
                    : label[l] < 6 ? bvh.data() + label[l] * plane
                                   : svh.data() + (label[l] - 6) * plane;
      insert(part, e, 0, dst);
    }
    st = corefl_state_upload(ctx, b, bvh.data(), r.n_scalar > 0 ? svh.data() : nullptr);
    if (st == 0 && r.tt) st = corefl_field_upload(ctx, b, COREFL_F_TVE, tveh.data());
    if (st != 0) return st;
  }
  if (solution_time) *solution_time = time;
  return 0;
}
