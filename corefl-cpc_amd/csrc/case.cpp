// Reference case directories: the grid, boundary and inner-face files the
// readGrid preprocessor writes under input/ (read as Mesh::read_grid,
// read_boundary and read_inner_interface do, src/Mesh.cpp:692-962), the
// ghost-node extension (Mesh::init_ghost_grid, src/Mesh.cpp:1013-1217), the
// metrics (Block::compute_jac_metric, src/Mesh.cpp:244-528), the initial
// state by groups (initialize_basic_variables, src/Initialize.cu:118-160,
// src/Field.cu:289-360) and the setup's boundary conditions, each state as
// the Inflow constructor derives it (src/BoundCond.cpp:150-330).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "corefl_internal.h"

using corefl::Setup;
using corefl::Status;

struct corefl_mesh {
  std::string dir, err;
  Setup setup;
  int ngg = 3, dim = 3;
  struct Blk {
    int mx = 0, my = 0, mz = 0;
    // coordinates over [-ngg-1, m+ngg] along each axis (i fastest)
    std::vector<double> x, y, z;
    // jac and the nine metrics over the padded extent [-ngg, m+ngg-1]
    std::vector<double> metric, jac;
    std::vector<corefl_boundary> faces;
    std::vector<corefl_inner_face> inner;
    long cx(int i, int j, int k, int g) const {
      const long nx = mx + 2L * (g + 1), ny = my + 2L * (g + 1);
      return (static_cast<long>(k) + g + 1) * nx * ny + (static_cast<long>(j) + g + 1) * nx + i + g + 1;
    }
  };
  std::vector<Blk> blocks;
};

namespace {

constexpr double kPi = 3.14159265358979323846;

int mfail(corefl_mesh *m, Status st, const std::string &msg) {
  m->err = msg;
  return static_cast<int>(st);
}

std::string rank_name(const std::string &prefix, int rank, const char *ext) {
  char buf[32];
  std::snprintf(buf, sizeof(buf), "%4d", rank);
  return prefix + buf + ext;
}

// Ghost layers from the inviscid and viscous schemes (src/Parameter.cpp:216-321).
int deduce_ngg(const Setup &s) {
  const int sch = corefl::get_or(s.ints, std::string("inviscid_scheme"), 3);
  const double thr = corefl::get_or(s.reals, std::string("shockSensor_threshold"), 0.0);
  const int recon = corefl::get_or(s.ints, std::string("reconstruction"), 2);
  int ngg = 2;
  if (sch == 51 || sch == 52) ngg = 3;
  if (sch == 71 || sch == 72) ngg = 4;
  if (sch >= 2 && sch <= 4 && (recon == 4 || recon == 5)) ngg = recon == 5 ? 4 : 3;
  (void)thr;
  const int vo = corefl::get_or(s.ints, std::string("viscous_order"), 2);
  if (vo == 8 && ngg < 4) ngg = 4;
  return ngg;
}

int read_grid(corefl_mesh *m, int rank) {
  const bool binary = corefl::get_or(m->setup.ints, std::string("gridIsBinary"), 1) != 0;
  const double scale = corefl::get_or(m->setup.reals, std::string("gridScale"), 1.0);
  const std::string path = rank_name(m->dir + "grid/grid", rank, binary ? ".dat" : ".grd");
  std::vector<std::vector<double>> xyz;
  std::vector<int> dims;
  int nb = 0;
  if (binary) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return mfail(m, Status::Io, "cannot open " + path);
    f.read(reinterpret_cast<char *>(&nb), sizeof(int));
    if (!f || nb < 1 || nb > 1 << 20) return mfail(m, Status::Parse, "bad block count in " + path);
    dims.resize(3 * static_cast<size_t>(nb));
    f.read(reinterpret_cast<char *>(dims.data()), 3 * nb * sizeof(int));
    if (!f) return mfail(m, Status::Parse, "truncated block sizes in " + path);
    for (int b = 0; b < nb; ++b) {
      const long n = static_cast<long>(dims[3 * b]) * dims[3 * b + 1] * dims[3 * b + 2];
      if (dims[3 * b] < 1 || dims[3 * b + 1] < 1 || dims[3 * b + 2] < 1 || n > (1L << 31))
        return mfail(m, Status::Parse, "bad block size in " + path);
      for (int a = 0; a < 3; ++a) {
        xyz.emplace_back(n);
        f.read(reinterpret_cast<char *>(xyz.back().data()), n * sizeof(double));
        if (!f) return mfail(m, Status::Parse, "truncated coordinates in " + path);
      }
    }
  } else {
    std::ifstream f(path);
    if (!f) return mfail(m, Status::Io, "cannot open " + path);
    std::string line;
    std::getline(f, line);
    nb = std::atoi(line.c_str());
    if (nb < 1 || nb > 1 << 20) return mfail(m, Status::Parse, "bad block count in " + path);
    dims.resize(3 * static_cast<size_t>(nb));
    for (int b = 0; b < nb; ++b) {
      std::getline(f, line);
      std::istringstream ls(line);
      ls >> dims[3 * b] >> dims[3 * b + 1] >> dims[3 * b + 2];
      if (dims[3 * b] < 1 || dims[3 * b + 1] < 1 || dims[3 * b + 2] < 1)
        return mfail(m, Status::Parse, "bad block size in " + path);
    }
    for (int b = 0; b < nb; ++b) {
      const long n = static_cast<long>(dims[3 * b]) * dims[3 * b + 1] * dims[3 * b + 2];
      for (int a = 0; a < 3; ++a) {
        xyz.emplace_back(n);
        for (long t = 0; t < n; ++t)
          if (!(f >> xyz.back()[t])) return mfail(m, Status::Parse, "truncated coordinates in " + path);
      }
    }
  }
  const int g = m->ngg;
  m->dim = dims[2] == 1 ? 2 : 3;
  m->blocks.resize(nb);
  for (int b = 0; b < nb; ++b) {
    auto &B = m->blocks[b];
    B.mx = dims[3 * b];
    B.my = dims[3 * b + 1];
    B.mz = dims[3 * b + 2];
    const long n = (B.mx + 2L * (g + 1)) * (B.my + 2L * (g + 1)) * (B.mz + 2L * (g + 1));
    B.x.assign(n, 0.0);
    B.y.assign(n, 0.0);
    B.z.assign(n, 0.0);
    for (int k = 0; k < B.mz; ++k)
      for (int j = 0; j < B.my; ++j)
        for (int i = 0; i < B.mx; ++i) {
          const long src = (static_cast<long>(k) * B.my + j) * B.mx + i, dst = B.cx(i, j, k, g);
          B.x[dst] = xyz[3 * b][src] * scale;
          B.y[dst] = xyz[3 * b + 1][src] * scale;
          B.z[dst] = xyz[3 * b + 2][src] * scale;
        }
  }
  return 0;
}

int read_boundary(corefl_mesh *m, int rank) {
  const std::string path = rank_name(m->dir + "boundary_condition/boundary", rank, ".txt");
  std::ifstream f(path);
  if (!f) return mfail(m, Status::Io, "cannot open " + path);
  for (auto &B : m->blocks) {
    int n = 0;
    if (!(f >> n) || n < 0) return mfail(m, Status::Parse, "bad face count in " + path);
    for (int q = 0; q < n; ++q) {
      int v[7];
      for (int &x : v)
        if (!(f >> x)) return mfail(m, Status::Parse, "truncated face line in " + path);
      corefl_boundary fb{};
      fb.type = v[6];
      fb.face = -1;
      fb.direction = 1;
      for (int a = 0; a < 3; ++a) {
        fb.range_start[a] = v[2 * a];
        fb.range_end[a] = v[2 * a + 1];
      }
      // Boundary::register_boundary (src/Mesh.cpp:14-32): the first axis whose range is one node is the
      // normal; index 0 there is the low face
      for (int a = 0; a < m->dim; ++a)
        if (fb.range_start[a] == fb.range_end[a]) {
          fb.face = a;
          if (fb.range_start[a] == 0) fb.direction = -1;
          break;
        }
      if (fb.face < 0) return mfail(m, Status::Parse, "a boundary face without a normal axis in " + path);
      for (int a = 0; a < 3; ++a)
        if (fb.range_end[a] < fb.range_start[a]) std::swap(fb.range_start[a], fb.range_end[a]);
      B.faces.push_back(fb);
    }
  }
  return 0;
}

int read_inner(corefl_mesh *m, int rank) {
  const std::string path = rank_name(m->dir + "boundary_condition/inner", rank, ".txt");
  std::ifstream f(path);
  if (!f) return 0;  // a single block without interfaces needs no file
  std::string line;
  std::getline(f, line);
  std::vector<int> counts(m->blocks.size());
  for (auto &c : counts)
    if (!(f >> c) || c < 0) return mfail(m, Status::Parse, "bad face counts in " + path);
  std::getline(f, line);
  std::getline(f, line);  // column header
  for (size_t b = 0; b < m->blocks.size(); ++b)
    for (int q = 0; q < counts[b]; ++q) {
      int s[7], t[7];
      for (int &x : s)
        if (!(f >> x)) return mfail(m, Status::Parse, "truncated inner face in " + path);
      for (int &x : t)
        if (!(f >> x)) return mfail(m, Status::Parse, "truncated inner face in " + path);
      corefl_inner_face fi{};
      std::copy(s, s + 6, fi.src);
      std::copy(t, t + 6, fi.tar);
      fi.target_block = t[6];
      if (t[6] < 1 || t[6] > static_cast<int>(m->blocks.size()))
        return mfail(m, Status::Parse, "inner face targets a block that does not exist in " + path);
      m->blocks[b].inner.push_back(fi);
    }
  return 0;
}

// Mirror extension of the nodes across every face, the interface ghost nodes from the neighbour, the
// edge and corner ghosts by the parallelogram rule (src/Mesh.cpp:1013-1217).
void extend_ghost_grid(corefl_mesh *m) {
  const int g = m->ngg, dim = m->dim;
  for (auto &B : m->blocks) {
    const int mx = B.mx, my = B.my, mz = B.mz;
    auto mirror = [&](long d, long o, long s) {
      B.x[d] = 2 * B.x[o] - B.x[s];
      B.y[d] = 2 * B.y[o] - B.y[s];
      B.z[d] = 2 * B.z[o] - B.z[s];
    };
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j) {
        for (int i = -1; i >= -g - 1; --i) mirror(B.cx(i, j, k, g), B.cx(0, j, k, g), B.cx(-i, j, k, g));
        for (int i = mx; i < mx + g + 1; ++i)
          mirror(B.cx(i, j, k, g), B.cx(mx - 1, j, k, g), B.cx(2 * (mx - 1) - i, j, k, g));
      }
    for (int k = 0; k < mz; ++k)
      for (int i = 0; i < mx; ++i) {
        for (int j = -1; j >= -g - 1; --j) mirror(B.cx(i, j, k, g), B.cx(i, 0, k, g), B.cx(i, -j, k, g));
        for (int j = my; j < my + g + 1; ++j)
          mirror(B.cx(i, j, k, g), B.cx(i, my - 1, k, g), B.cx(i, 2 * (my - 1) - j, k, g));
      }
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        if (dim == 3) {
          for (int k = -1; k >= -g - 1; --k) mirror(B.cx(i, j, k, g), B.cx(i, j, 0, g), B.cx(i, j, -k, g));
          for (int k = mz; k < mz + g + 1; ++k)
            mirror(B.cx(i, j, k, g), B.cx(i, j, mz - 1, g), B.cx(i, j, 2 * (mz - 1) - k, g));
        } else {
          // 2-D: unit spacing in z (src/Mesh.cpp:1063-1079)
          for (int k = -1; k >= -g - 1; --k) {
            const long d = B.cx(i, j, k, g), o = B.cx(i, j, 0, g);
            B.x[d] = B.x[o];
            B.y[d] = B.y[o];
            B.z[d] = B.z[o] + k;
          }
          for (int k = mz; k < mz + g + 1; ++k) {
            const long d = B.cx(i, j, k, g), o = B.cx(i, j, mz - 1, g);
            B.x[d] = B.x[o];
            B.y[d] = B.y[o];
            B.z[d] = B.z[o] + k - mz + 1;
          }
        }
      }
  }
  // interface ghosts from the neighbour's nodes (init_inner_ghost_grid, src/Mesh.cpp:1169-1217)
  for (auto &B : m->blocks) {
    for (const auto &fi : B.inner) {
      corefl::InnerDev f{};
      corefl::register_inner(fi, dim, g, f);
      const auto &N = m->blocks[f.target_block];
      int lo[3], hi[3], tst[3];
      for (int a = 0; a < 3; ++a) {
        lo[a] = f.range_start[a];
        hi[a] = f.range_end[a];
        tst[a] = f.target_start[a];
      }
      // the registered loop covers the face node and its ghost layers; this copy shifts it one node
      // outward, as the reference does for the ghost-grid assignment
      const int shift = f.direction < 0 ? -1 : 1;
      lo[f.face] += shift;
      hi[f.face] += shift;
      tst[f.target_face] += f.target_direction > 0 ? -1 : 1;
      int d[3];
      for (int i1 = lo[0]; f.loop_dir[0] * (i1 - hi[0]) != 1; i1 += f.loop_dir[0]) {
        d[0] = i1 - lo[0];
        for (int j1 = lo[1]; f.loop_dir[1] * (j1 - hi[1]) != 1; j1 += f.loop_dir[1]) {
          d[1] = j1 - lo[1];
          for (int k1 = lo[2]; f.loop_dir[2] * (k1 - hi[2]) != 1; k1 += f.loop_dir[2]) {
            d[2] = k1 - lo[2];
            const int i2 = tst[0] + f.target_loop_dir[0] * d[f.src_tar[0]];
            const int j2 = tst[1] + f.target_loop_dir[1] * d[f.src_tar[1]];
            const int k2 = tst[2] + f.target_loop_dir[2] * d[f.src_tar[2]];
            if (i1 < -g - 1 || i1 > B.mx + g || j1 < -g - 1 || j1 > B.my + g || k1 < -g - 1 || k1 > B.mz + g ||
                i2 < -g - 1 || i2 > N.mx + g || j2 < -g - 1 || j2 > N.my + g || k2 < -g - 1 || k2 > N.mz + g)
              continue;
            const long s = N.cx(i2, j2, k2, g), t = B.cx(i1, j1, k1, g);
            B.x[t] = N.x[s];
            B.y[t] = N.y[s];
            B.z[t] = N.z[s];
          }
        }
      }
    }
  }
  // edges and corners (parallelogram rule)
  for (auto &B : m->blocks) {
    const int mx = B.mx, my = B.my, mz = B.mz;
    auto para = [&](long d, long a, long b, long o) {
      B.x[d] = B.x[a] + B.x[b] - B.x[o];
      B.y[d] = B.y[a] + B.y[b] - B.y[o];
      B.z[d] = B.z[a] + B.z[b] - B.z[o];
    };
    for (int i = 1; i <= g + 1; ++i)
      for (int j = 1; j <= g + 1; ++j)
        for (int k = 0; k < mz; ++k) {
          const int i1 = -i, j1 = -j, i2 = mx - 1 + i, j2 = my - 1 + j;
          para(B.cx(i1, j1, k, g), B.cx(i1, 0, k, g), B.cx(0, j1, k, g), B.cx(0, 0, k, g));
          para(B.cx(i2, j2, k, g), B.cx(i2, my - 1, k, g), B.cx(mx - 1, j2, k, g), B.cx(mx - 1, my - 1, k, g));
          para(B.cx(i1, j2, k, g), B.cx(i1, my - 1, k, g), B.cx(0, j2, k, g), B.cx(0, my - 1, k, g));
          para(B.cx(i2, j1, k, g), B.cx(i2, 0, k, g), B.cx(mx - 1, j1, k, g), B.cx(mx - 1, 0, k, g));
        }
    for (int i = 1; i <= g + 1; ++i)
      for (int j = 0; j < my; ++j)
        for (int k = 1; k <= g + 1; ++k) {
          const int i1 = -i, k1 = -k, i2 = mx - 1 + i, k2 = mz - 1 + k;
          para(B.cx(i1, j, k1, g), B.cx(i1, j, 0, g), B.cx(0, j, k1, g), B.cx(0, j, 0, g));
          para(B.cx(i2, j, k2, g), B.cx(i2, j, mz - 1, g), B.cx(mx - 1, j, k2, g), B.cx(mx - 1, j, mz - 1, g));
          para(B.cx(i1, j, k2, g), B.cx(i1, j, mz - 1, g), B.cx(0, j, k2, g), B.cx(0, j, mz - 1, g));
          para(B.cx(i2, j, k1, g), B.cx(i2, j, 0, g), B.cx(mx - 1, j, k1, g), B.cx(mx - 1, j, 0, g));
        }
    for (int i = -g - 1; i <= mx + g; ++i)
      for (int j = 1; j <= g + 1; ++j)
        for (int k = 1; k <= g + 1; ++k) {
          const int j1 = -j, k1 = -k, j2 = my - 1 + j, k2 = mz - 1 + k;
          para(B.cx(i, j1, k1, g), B.cx(i, j1, 0, g), B.cx(i, 0, k1, g), B.cx(i, 0, 0, g));
          para(B.cx(i, j2, k2, g), B.cx(i, j2, mz - 1, g), B.cx(i, my - 1, k2, g), B.cx(i, my - 1, mz - 1, g));
          para(B.cx(i, j1, k2, g), B.cx(i, j1, mz - 1, g), B.cx(i, 0, k2, g), B.cx(i, 0, mz - 1, g));
          para(B.cx(i, j2, k1, g), B.cx(i, j2, 0, g), B.cx(i, my - 1, k1, g), B.cx(i, my - 1, 0, g));
        }
  }
}

// Central-difference Jacobian and metrics over the padded extent; a node whose Jacobian is not positive
// takes the values of its neighbour toward the interior (src/Mesh.cpp:244-528).
int compute_metrics(corefl_mesh *m) {
  const int g = m->ngg;
  for (size_t b = 0; b < m->blocks.size(); ++b) {
    auto &B = m->blocks[b];
    const long nx = B.mx + 2L * g, ny = B.my + 2L * g, nz = B.mz + 2L * g, sz = nx * ny * nz;
    B.metric.assign(9 * sz, 0.0);
    B.jac.assign(sz, 0.0);
    const long si = 1, sj = B.mx + 2L * (g + 1), sk = sj * (B.my + 2L * (g + 1));
    // interior first, then outward layer by layer so a fallback finds its neighbour computed
    std::vector<std::array<int, 3>> order;
    order.reserve(sz);
    for (int k = -g; k < B.mz + g; ++k)
      for (int j = -g; j < B.my + g; ++j)
        for (int i = -g; i < B.mx + g; ++i) order.push_back({i, j, k});
    auto depth = [&](const std::array<int, 3> &p) {
      auto out = [](int v, int n) { return v < 0 ? -v : (v >= n ? v - n + 1 : 0); };
      return std::max({out(p[0], B.mx), out(p[1], B.my), out(p[2], B.mz)});
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](const auto &a, const auto &c) { return depth(a) < depth(c); });
    for (const auto &p : order) {
      const int i = p[0], j = p[1], k = p[2];
      const long c = B.cx(i, j, k, g);
      const double dxd1 = (B.x[c + si] - B.x[c - si]) * 0.5, dyd1 = (B.y[c + si] - B.y[c - si]) * 0.5,
                   dzd1 = (B.z[c + si] - B.z[c - si]) * 0.5;
      const double dxd2 = (B.x[c + sj] - B.x[c - sj]) * 0.5, dyd2 = (B.y[c + sj] - B.y[c - sj]) * 0.5,
                   dzd2 = (B.z[c + sj] - B.z[c - sj]) * 0.5;
      const double dxd3 = (B.x[c + sk] - B.x[c - sk]) * 0.5, dyd3 = (B.y[c + sk] - B.y[c - sk]) * 0.5,
                   dzd3 = (B.z[c + sk] - B.z[c - sk]) * 0.5;
      const double jac = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                         dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
      const long q = (static_cast<long>(k) + g) * nx * ny + (static_cast<long>(j) + g) * nx + i + g;
      if (!(jac > 0)) {
        if (depth(p) == 0) {
          char buf[160];
          std::snprintf(buf, sizeof(buf), "non-positive Jacobian at (%d, %d, %d) of block %zu", i, j, k, b);
          return mfail(m, Status::Parse, buf);
        }
        // neighbour one node toward the interior along the axis that is farthest out
        int ii = i, jj = j, kk = k;
        auto out = [](int v, int n) { return v < 0 ? -v : (v >= n ? v - n + 1 : 0); };
        const int oi = out(i, B.mx), oj = out(j, B.my), ok = out(k, B.mz);
        if (ok >= oi && ok >= oj)
          kk += k < 0 ? 1 : -1;
        else if (oj >= oi)
          jj += j < 0 ? 1 : -1;
        else
          ii += i < 0 ? 1 : -1;
        const long qn = (static_cast<long>(kk) + g) * nx * ny + (static_cast<long>(jj) + g) * nx + ii + g;
        B.jac[q] = B.jac[qn];
        for (int l = 0; l < 9; ++l) B.metric[q + l * sz] = B.metric[qn + l * sz];
        continue;
      }
      B.jac[q] = jac;
      B.metric[q] = (dyd2 * dzd3 - dyd3 * dzd2) / jac;
      B.metric[q + sz] = (dxd3 * dzd2 - dxd2 * dzd3) / jac;
      B.metric[q + 2 * sz] = (dxd2 * dyd3 - dxd3 * dyd2) / jac;
      B.metric[q + 3 * sz] = (dyd3 * dzd1 - dyd1 * dzd3) / jac;
      B.metric[q + 4 * sz] = (dxd1 * dzd3 - dxd3 * dzd1) / jac;
      B.metric[q + 5 * sz] = (dxd3 * dyd1 - dxd1 * dyd3) / jac;
      B.metric[q + 6 * sz] = (dyd1 * dzd2 - dyd2 * dzd1) / jac;
      B.metric[q + 7 * sz] = (dxd2 * dzd1 - dxd1 * dzd2) / jac;
      B.metric[q + 8 * sz] = (dxd1 * dyd2 - dxd2 * dyd1) / jac;
    }
  }
  return 0;
}

// The state a setup struct names, as the Inflow constructor derives it (src/BoundCond.cpp:150-330):
// species from the keys named after them, T from (p, rho) when not given, the velocity magnitude from
// the Mach number, direction (u, v, w) or angle of attack, rho from the equation of state.
struct State {
  double rho, u, v, w, p, T, tve;
  double sv[COREFL_MAX_SPEC + 1];
  double mw;
};

int inflow_state(corefl_mesh *m, const std::string &name, const corefl_chem_tables *t, State &st) {
  auto it = m->setup.structs.find(name);
  if (it == m->setup.structs.end()) return mfail(m, Status::Parse, "the setup has no struct " + name);
  const auto &s = it->second;
  double mach = corefl::struct_real(s, "mach", -1), p = corefl::struct_real(s, "pressure", -1);
  double T = corefl::struct_real(s, "temperature", -1), tve = corefl::struct_real(s, "Tve", -1);
  double vel = corefl::struct_real(s, "velocity", 0), rho = corefl::struct_real(s, "density", -1);
  const double aoa = corefl::struct_real(s, "aoa", -1000);
  double u = corefl::struct_real(s, "u", 1), v = corefl::struct_real(s, "v", 0), w = corefl::struct_real(s, "w", 0);
  std::memset(&st, 0, sizeof(st));
  double mw = 28.9647;
  const int ns = t ? t->n_spec : 0;
  if (ns > 0) {
    double imw = 0;
    for (int l = 0; l < ns; ++l) {
      st.sv[l] = corefl::struct_real(s, corefl::species_name(*t, m->dir, m->setup, l), 0.0);
      imw += st.sv[l] / t->mw[l];
    }
    mw = 1 / imw;
  }
  if (T < 0) T = p * mw / (rho * corefl::kRu);
  if (tve < 0) tve = T;
  // gamma and the sound speed at T (translational-rotational and vibrational-electronic together, as
  // compute_boundary_gamma_and_sound_speed does for an equilibrium state)
  double gamma = 1.4;
  const double R = corefl::kRu / mw;
  if (ns > 0) {
    double cp = 0;
    for (int l = 0; l < ns; ++l) {
      double cpl = 0;
      if (t->nasa7or9 == 9) {
        const int n = t->n_range[l];
        const double tt = std::min(std::max(T, t->t_cut[l][0]), t->t_cut[l][n]);
        int j = 0;
        while (j < n - 1 && t->t_cut[l][j + 1] < tt) ++j;
        const double *c = t->poly9[l][j];
        cpl = (c[0] / (tt * tt) + c[1] / tt + c[2] + c[3] * tt + c[4] * tt * tt + c[5] * tt * tt * tt +
               c[6] * tt * tt * tt * tt) * t->gas_const[l];
      } else {
        const double tt = std::min(std::max(T, t->t_low[l]), t->t_high[l]);
        const double *c = tt < t->t_mid[l] ? t->low_coeff[l] : t->high_coeff[l];
        cpl = (c[0] + c[1] * tt + c[2] * tt * tt + c[3] * tt * tt * tt + c[4] * tt * tt * tt * tt) * t->gas_const[l];
      }
      cp += st.sv[l] * cpl;
    }
    gamma = cp / (cp - R);
  }
  const double c = std::sqrt(gamma * R * T);
  if (mach < 0)
    mach = vel / c;
  else
    vel = mach * c;
  if (aoa > -1000) {
    u = std::cos(aoa * kPi / 180) * vel;
    v = std::sin(aoa * kPi / 180) * vel;
    w = 0;
  } else {
    u *= vel;
    v *= vel;
    w *= vel;
  }
  if (rho < 0) {
    if (p < 0) return mfail(m, Status::Parse, "struct " + name + " gives neither density nor pressure");
    rho = p * mw / (corefl::kRu * T);
  }
  if (p < 0) p = rho * T * corefl::kRu / mw;
  st.rho = rho;
  st.u = u;
  st.v = v;
  st.w = w;
  st.p = p;
  st.T = T;
  st.tve = tve;
  st.mw = mw;
  return 0;
}

}  // namespace

extern "C" {

int corefl_mesh_load(const char *case_dir, int rank, corefl_mesh **out) {
  if (!case_dir || !out) return static_cast<int>(Status::Arg);
  auto *m = new corefl_mesh();
  *out = m;
  m->dir = std::string(case_dir) + "/input/";
  std::string err;
  if (!corefl::read_setup(m->dir + "setup.txt", m->setup, err)) return mfail(m, Status::Io, err);
  m->ngg = deduce_ngg(m->setup);
  int st;
  if ((st = read_grid(m, rank))) return st;
  if ((st = read_boundary(m, rank))) return st;
  if ((st = read_inner(m, rank))) return st;
  for (auto &B : m->blocks)
    for (const auto &fi : B.inner) {
      const auto &T = m->blocks[fi.target_block - 1];
      // both ranges inside their blocks (1-based, signs give the running direction)
      for (int a = 0; a < 3; ++a) {
        const int ms[3] = {B.mx, B.my, B.mz}, mt[3] = {T.mx, T.my, T.mz};
        for (int e = 0; e < 2; ++e) {
          if (std::abs(fi.src[2 * a + e]) > ms[a] || std::abs(fi.tar[2 * a + e]) > mt[a] ||
              fi.src[2 * a + e] == 0 || fi.tar[2 * a + e] == 0)
            return mfail(m, Status::Parse, "an inner face range lies outside its block");
        }
      }
    }
  extend_ghost_grid(m);
  return compute_metrics(m);
}

int corefl_mesh_n_block(const corefl_mesh *m) { return m ? static_cast<int>(m->blocks.size()) : 0; }

int corefl_mesh_block(const corefl_mesh *m, int blk, corefl_block_desc *desc) {
  if (!m || !desc || blk < 0 || blk >= static_cast<int>(m->blocks.size())) return static_cast<int>(Status::Arg);
  const auto &B = m->blocks[blk];
  desc->mx = B.mx;
  desc->my = B.my;
  desc->mz = B.mz;
  desc->ngg = m->ngg;
  desc->metric = B.metric.data();
  desc->jac = B.jac.data();
  desc->n_boundary = static_cast<int>(B.faces.size());
  desc->boundary = B.faces.data();
  desc->n_inner = static_cast<int>(B.inner.size());
  desc->inner = B.inner.data();
  return 0;
}

int corefl_mesh_coords(const corefl_mesh *m, int blk, double *xyz) {
  if (!m || !xyz || blk < 0 || blk >= static_cast<int>(m->blocks.size())) return static_cast<int>(Status::Arg);
  const auto &B = m->blocks[blk];
  const int g = m->ngg;
  const long nx = B.mx + 2L * g, ny = B.my + 2L * g, sz = nx * ny * (B.mz + 2L * g);
  for (int k = -g; k < B.mz + g; ++k)
    for (int j = -g; j < B.my + g; ++j)
      for (int i = -g; i < B.mx + g; ++i) {
        const long q = (static_cast<long>(k) + g) * nx * ny + (static_cast<long>(j) + g) * nx + i + g;
        const long c = B.cx(i, j, k, g);
        xyz[q] = B.x[c];
        xyz[q + sz] = B.y[c];
        xyz[q + 2 * sz] = B.z[c];
      }
  return 0;
}

int corefl_mesh_initial_state(corefl_mesh *m, int blk, const corefl_chem_tables *tables, double *bv, double *sv,
                              double *tve) {
  if (!m || !bv || blk < 0 || blk >= static_cast<int>(m->blocks.size())) return static_cast<int>(Status::Arg);
  const int ns = tables ? tables->n_spec : 0;
  if (ns > 0 && !sv) return static_cast<int>(Status::Arg);
  const Setup &s = m->setup;
  std::vector<State> states(1);
  std::vector<double> box;
  int st;
  const std::string def = corefl::get_or(s.strings, std::string("default_init"), std::string(""));
  if ((st = inflow_state(m, def, tables, states[0]))) return st;
  const int groups = corefl::get_or(s.ints, std::string("groups_init"), 1);
  for (int l = 0; l < groups - 1; ++l) {
    const std::string pn = "init_cond_" + std::to_string(l);
    auto it = s.structs.find(pn);
    if (it == s.structs.end()) return mfail(m, Status::Parse, "the setup has no struct " + pn);
    for (const char *k : {"x0", "x1", "y0", "y1", "z0", "z1"}) box.push_back(corefl::struct_real(it->second, k, 0));
    std::string name = pn;
    auto nit = it->second.find("name");
    if (nit != it->second.end())
      if (auto p = std::get_if<std::string>(&nit->second)) name = *p;
    states.emplace_back();
    if ((st = inflow_state(m, name, tables, states.back()))) return st;
  }
  const auto &B = m->blocks[blk];
  const int g = m->ngg;
  const long nx = B.mx + 2L * g, ny = B.my + 2L * g, sz = nx * ny * (B.mz + 2L * g);
  for (int k = -g; k < B.mz + g; ++k)
    for (int j = -g; j < B.my + g; ++j)
      for (int i = -g; i < B.mx + g; ++i) {
        const long q = (static_cast<long>(k) + g) * nx * ny + (static_cast<long>(j) + g) * nx + i + g;
        const long c = B.cx(i, j, k, g);
        int pick = 0;
        for (int l = 0; l < groups - 1; ++l) {
          const double *r = &box[6 * l];
          if (B.x[c] >= r[0] && B.x[c] <= r[1] && B.y[c] >= r[2] && B.y[c] <= r[3] && B.z[c] >= r[4] &&
              B.z[c] <= r[5]) {
            pick = l + 1;
            break;
          }
        }
        const State &S = states[pick];
        bv[q] = S.rho;
        bv[q + sz] = S.u;
        bv[q + 2 * sz] = S.v;
        bv[q + 3 * sz] = S.w;
        bv[q + 4 * sz] = S.p;
        bv[q + 5 * sz] = S.T;
        for (int l = 0; l < ns; ++l) sv[q + l * sz] = S.sv[l];
        if (tve) tve[q] = S.tve;
      }
  return 0;
}

int corefl_mesh_conditions(corefl_mesh *m, const corefl_chem_tables *tables, corefl_bc_condition *conds, int max) {
  if (!m || max < 0 || (max > 0 && !conds)) return static_cast<int>(Status::Arg);
  const Setup &s = m->setup;
  auto it = s.string_arrays.find("boundary_conditions");
  if (it == s.string_arrays.end()) return 0;
  int n = 0;
  for (const auto &name : it->second) {
    auto sit = s.structs.find(name);
    if (sit == s.structs.end()) return mfail(m, Status::Parse, "the setup has no struct " + name);
    const auto &S = sit->second;
    std::string type;
    if (auto p = S.find("type"); p != S.end())
      if (auto q = std::get_if<std::string>(&p->second)) type = *q;
    corefl_bc_condition c{};
    c.label = static_cast<int>(corefl::struct_real(S, "label", 0));
    if (type == "wall") {
      c.kind = COREFL_BC_WALL;
      std::string tt;
      if (auto p = S.find("thermal_type"); p != S.end())
        if (auto q = std::get_if<std::string>(&p->second)) tt = *q;
      c.thermal_type = tt == "adiabatic" ? 1 : 0;
      c.wall_temperature = corefl::struct_real(S, "temperature", 300.0);
      c.temperature_ve = corefl::struct_real(S, "Tve", -1.0);
    } else if (type == "inflow") {
      c.kind = COREFL_BC_INFLOW;
      State st;
      int e;
      if ((e = inflow_state(m, name, tables, st))) return e;
      c.inflow_sub = static_cast<int>(corefl::struct_real(S, "inflow_sub", 0));
      c.density = st.rho;
      c.u = st.u;
      c.v = st.v;
      c.w = st.w;
      c.pressure = st.p;
      c.temperature = st.T;
      c.mw = st.mw;
      for (int l = 0; l < (tables ? tables->n_spec : 0); ++l) c.sv[l] = st.sv[l];
      c.temperature_ve = st.tve;
    } else if (type == "outflow") {
      c.kind = COREFL_BC_OUTFLOW;
    } else if (type == "symmetry") {
      c.kind = COREFL_BC_SYMMETRY;
    } else if (type == "periodic") {
      c.kind = COREFL_BC_PERIODIC;
    } else {
      return mfail(m, Status::Parse, "boundary type " + type + " of " + name + " is not supported");
    }
    if (n < max) conds[n] = c;
    ++n;
  }
  return n;
}

const char *corefl_mesh_error(const corefl_mesh *m) { return m ? m->err.c_str() : "null mesh"; }

void corefl_mesh_free(corefl_mesh *m) { delete m; }

}  // extern "C"
