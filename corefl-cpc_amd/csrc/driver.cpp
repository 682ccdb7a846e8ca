// Host driver behind the C ABI: owns device memory for every block, runs the
// per-stage operators in the order RK3<mix_model> runs them
// (src/RK.cuh:107-272) and initialisation in the order of
// Driver::initialize_computation (src/Driver.cu:126-211).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "corefl_internal.h"

struct corefl_ctx {
  corefl::Run run{};
  corefl_run_desc desc{};
  corefl_chem_tables host_tables{};
  corefl::Tab *d_tables = nullptr;
  corefl::ChemJit chem_jit;  // mechanism-specialised chemistry kernel (chem_kernel 0)
  double *d_scalar = nullptr;  // device scalars: [0] block dt minimum
  // state flag of the update kernels (Zone::bad), its pinned host copy and the event that tells the copy landed
  int *d_bad = nullptr, *h_bad = nullptr;
  hipEvent_t bad_ev = nullptr;
  bool bad_posted = false;
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  std::vector<corefl_bc_condition> conditions;  // registered per label (corefl_bc_set)
  struct Face {
    corefl_boundary b;
    int kind;                    // COREFL_BC_*
    corefl_bc_condition cond;
  };
  struct Block {
    corefl::Zone z{};
    std::vector<Face> faces;
    std::vector<corefl::InnerDev> inner;
    std::vector<void *> allocs;
    int shape[3] = {0, 0, 0};  // tile shape of each direction's characteristic WENO sweep
  };
  std::vector<Block> blocks;
  bool inner_checked = false;  // inner-face targets validated against the blocks present
  // profiling
  bool profiling = false;
  struct Ev {
    int op;
    hipEvent_t a, b;
  };
  std::vector<Ev> events;
  std::vector<hipEvent_t> event_pool;
  size_t pool_used = 0;
  // captured step (desc.step_graph): one stream per block joined to `stream` by events while the step
  // is recorded, the blocks' dt minima and the instantiated graph
  std::vector<hipStream_t> bstreams;
  std::vector<hipEvent_t> join_ev;
  hipEvent_t fork_ev = nullptr;
  hipStream_t cur = nullptr;  // stream op_block launches on (nullptr: `stream`)
  bool qn_in_update = false;  // recording a step whose first-stage update saves qn (no separate copy)
  // multi-block launches of a recorded multi-block step: the zones and the outflow faces in device memory
  bool batched = false;
  corefl::Zone *d_zones = nullptr;
  corefl::BcItem *d_outflow = nullptr;
  int n_outflow = 0;
  long mb_interior = 0, mb_padded = 0, mb_face = 0;  // largest block / face
  double *d_bmin = nullptr;   // [n_block] block dt minima
  hipGraphExec_t graph_exec = nullptr;
  double graph_dt = 0.0;      // fixed dt recorded in the graph
  size_t graph_nb = 0;        // blocks d_bmin holds
  bool capturing = false;
};

namespace corefl {
namespace {

// E_ve of species l at t on the host, Species::compute_ve_energy (src/ChemData.cpp:238-261)
double host_ve_energy(const corefl_chem_tables &t, int l, double temp_in) {
  if (!t.two_temperature) return 0.0;
  const double temp = std::max(temp_in, 1.0);
  const double r = kRu / t.mw[l];
  double e = 0.0;
  if (t.theta_v[l] > 0) {
    const double x = t.theta_v[l] / temp;
    if (x < 200) e += r * t.theta_v[l] / std::expm1(x);
  }
  if (t.n_elec[l] > 0) {
    double z = 0.0, w = 0.0;
    for (int j = 0; j < t.n_elec[l]; ++j) {
      const double b = t.elec_g[l][j] * std::exp(-t.elec_theta[l][j] / temp);
      z += b;
      w += b * t.elec_theta[l][j];
    }
    if (z > 0) e += r * w / z;
  }
  return e;
}

// share of the step's dt in RK stage rk (the SSP-RK3 c coefficients), src/RK.cuh:91
constexpr double kStageDtScale[3] = {1.0, 0.25, 2.0 / 3.0};

int fail(corefl_ctx *c, Status s, const std::string &msg) {
  if (c) c->err = msg;
  return static_cast<int>(s);
}

int hip_fail(corefl_ctx *c, hipError_t e, const char *where) {
  return fail(c, Status::Hip, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                 \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

hipEvent_t pool_event(corefl_ctx *c) {
  if (c->pool_used == c->event_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->event_pool.push_back(e);
  }
  return c->event_pool[c->pool_used++];
}

struct OpTimer {
  corefl_ctx *c;
  int op;
  hipEvent_t a = nullptr;
  OpTimer(corefl_ctx *ctx, int o) : c(ctx), op(o) {
    if (c->profiling && !c->capturing) {
      a = pool_event(c);
      if (a) (void)hipEventRecord(a, c->stream);
    }
  }
  ~OpTimer() {
    if (c->profiling && !c->capturing && a) {
      hipEvent_t b = pool_event(c);
      if (b) {
        (void)hipEventRecord(b, c->stream);
        c->events.push_back({op, a, b});
      }
    }
  }
};

long field_planes(const corefl_ctx *c, int f) {
  const Run &r = c->run;
  switch (f) {
    case COREFL_F_CV:
    case COREFL_F_QN:
    case COREFL_F_DQ:
    case COREFL_F_FFLUX:
    case COREFL_F_GFLUX:
    case COREFL_F_HFLUX:
      return r.nv;
    case COREFL_F_BV:
      return 6;
    case COREFL_F_SV:
      return r.n_scalar > 0 ? r.n_scalar : 1;
    case COREFL_F_METRIC:
      return 9;
    case COREFL_F_RHOD:
      return r.ns > 0 ? r.ns : 1;
    default:
      return 1;
  }
}

double *field_ptr(const Zone &z, int f) {
  switch (f) {
    case COREFL_F_CV: return z.cv;
    case COREFL_F_BV: return z.bv;
    case COREFL_F_SV: return z.sv;
    case COREFL_F_QN: return z.qn;
    case COREFL_F_DQ: return z.dq;
    case COREFL_F_FFLUX: return z.ff;
    case COREFL_F_GFLUX: return z.gf;
    case COREFL_F_HFLUX: return z.hf;
    case COREFL_F_METRIC: return z.metric;
    case COREFL_F_JAC: return z.jac;
    case COREFL_F_MUL: return z.mul;
    case COREFL_F_KAPPA: return z.kappa;
    case COREFL_F_GAMMA: return z.gam;
    case COREFL_F_ACOUSTIC: return z.acoustic;
    case COREFL_F_MACH: return z.mach;
    case COREFL_F_RHOD: return z.rhoD;
    case COREFL_F_RTS: return z.rts;
    case COREFL_F_DT_LOCAL: return z.dt_local;
    case COREFL_F_TVE: return z.tve;
    case COREFL_F_KAPPA_VE: return z.kappa_ve;
    case COREFL_F_SHOCK_SENSOR: return z.sensor;
    default: return nullptr;
  }
}

bool interior_field(int f) { return f == COREFL_F_DQ || f == COREFL_F_DT_LOCAL; }

long field_doubles(const corefl_ctx *c, const Zone &z, int f) {
  return field_planes(c, f) * (interior_field(f) ? z.dsz : z.sz);
}

// The block's periodic faces in block order, each launched by `one(face)`, except that (with `pairs`) a face
// followed (among the periodic faces) by the opposite face of the same axis over the same tangential ranges goes
// with it to `pair(low face)` as one launch, when the block holds at least 2 ngg + 1 nodes along the axis (so no
// ghost copy reads a node the averaging writes).  The launches keep their order.  The pair relies on the second
// face's averaging of the coincident nodes changing nothing: true of the viscous fluxes and of a
// one-temperature state, not of a two-temperature one (the first face's conservative update re-derives E_ve).
template <class One, class Pair>
hipError_t periodic_launches(const corefl_ctx::Block &b, bool pairs, One one, Pair pair) {
  std::vector<const corefl_boundary *> per;
  for (const auto &f : b.faces)
    if (f.kind == COREFL_BC_PERIODIC) per.push_back(&f.b);
  for (size_t q = 0; q < per.size(); ++q) {
    const corefl_boundary &x = *per[q];
    if (pairs && q + 1 < per.size()) {
      const corefl_boundary &y = *per[q + 1];
      const int a = x.face, n = a == 0 ? b.z.mx : (a == 1 ? b.z.my : b.z.mz);
      bool same = y.face == a && y.direction == -x.direction && n >= 2 * b.z.ngg + 1;
      for (int t = 0; t < 3 && same; ++t)
        if (t != a) same = x.range_start[t] == y.range_start[t] && x.range_end[t] == y.range_end[t];
      if (same) {
        if (hipError_t e = pair(x.direction < 0 ? x : y); e != hipSuccess) return e;
        ++q;
        continue;
      }
    }
    if (hipError_t e = one(x); e != hipSuccess) return e;
  }
  return hipSuccess;
}

// ---- operators on one block ----
int op_block(corefl_ctx *c, int op, corefl_ctx::Block &b, int rk, double dt) {
  const Zone &z = b.z;
  const Run &r = c->run;
  hipStream_t s = c->cur ? c->cur : c->stream;
  // a recorded step reads dt from device memory (d_scalar[1])
  const double *dtp = c->capturing ? c->d_scalar + 1 : nullptr;
  OpTimer timer(c, op);
  switch (op) {
    case COREFL_OP_ZERO_DQ:
      HIP_TRY(c, launch_fill(z.dq, z.dsz * r.nv, 0.0, s));
      break;
    case COREFL_OP_VISCOUS:
    case COREFL_OP_HALO:
      return fail(c, Status::Arg, "this operator couples the blocks and runs on all of them");
    case COREFL_OP_CHEMISTRY:
      if (r.mixture && r.reaction) {
        if (c->chem_jit.fn)
          HIP_TRY(c, chem_jit_launch(c->chem_jit, z, s));
        else
          HIP_TRY(c, launch_chemistry(z, r, c->d_tables, s));
      }
      break;
    case COREFL_OP_INVISCID_X:
    case COREFL_OP_INVISCID_Y:
    case COREFL_OP_INVISCID_Z: {
      const int dir = op - COREFL_OP_INVISCID_X;
      if (dir == 2 && z.mz == 1) break;
      if (r.scheme == 51 || r.scheme == 52 || r.scheme == 71 || r.scheme == 72)
        HIP_TRY(c, launch_inviscid(z, r, c->d_tables, dir, s, b.shape[dir]));
      else
        HIP_TRY(c, launch_inviscid_pv(z, r, c->d_tables, dir, s));
      break;
    }
    case COREFL_OP_UPDATE_RK:
      if (rk < 0 || rk > 2) return fail(c, Status::Arg, "rk stage must be 0, 1 or 2");
      HIP_TRY(c, launch_update_rk(z, r, c->d_tables, dt, rk, s, dtp, nullptr, c->qn_in_update && rk == 0));
      // apply_vt_relaxation_rk with the stage's dt share {1, 1/4, 2/3}, src/RK.cuh:91, 200-203
      if (r.tt && r.vt)
        HIP_TRY(c, launch_vt_relaxation(z, r, c->d_tables, kStageDtScale[rk] * dt, s, dtp, kStageDtScale[rk]));
      break;
    case COREFL_OP_BC:
      // apply_boundary_conditions in its priority order (src/BoundCond.cu:3846-3997):
      // outflow, inflow, symmetry, wall, periodic; within a kind, faces in block order
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_OUTFLOW) HIP_TRY(c, launch_bc_outflow(z, r, c->d_tables, f.b, s));
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_INFLOW) HIP_TRY(c, launch_bc_inflow(z, r, c->d_tables, f.b, f.cond, s));
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_SYMMETRY) HIP_TRY(c, launch_bc_symmetry(z, r, c->d_tables, f.b, s));
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_WALL) HIP_TRY(c, launch_bc_wall(z, r, c->d_tables, f.b, f.cond, s));
      HIP_TRY(c, periodic_launches(
                     b, !r.tt, [&](const corefl_boundary &f) { return launch_bc_periodic(z, r, c->d_tables, f, s); },
                     [&](const corefl_boundary &lo) { return launch_bc_periodic_pair(z, r, c->d_tables, lo, s); }));
      break;
    case COREFL_OP_LIMIT_FLOW:
      HIP_TRY(c, launch_limit_flow(z, r, c->d_tables, s));
      break;
    case COREFL_OP_SHOCK_SENSOR:
      if (!r.hybrid) return fail(c, Status::Arg, "the shock sensor needs a hybrid WENO run (shockSensor_threshold > 1e-10)");
      HIP_TRY(c, launch_shock_sensor(z, r, s));
      break;
    case COREFL_OP_K_GRADIENT:
      if (z.mz == 1) HIP_TRY(c, launch_k_gradient(z, r, s));
      break;
    case COREFL_OP_PROPERTIES:
      HIP_TRY(c, launch_properties(z, r, c->d_tables, s));
      break;
    case COREFL_OP_CV_FROM_BV:
      HIP_TRY(c, launch_cv_from_bv(z, r, c->d_tables, s));
      break;
    case COREFL_OP_LOCAL_DT: {
      // a recorded step keeps each block's minimum in its own slot (the blocks run concurrently); Wu's
      // splitting takes local_time_step_without_reaction (src/TimeAdvanceFunc.cu:21-112)
      Run rr = r;
      if (r.wu) rr.reaction = 0;
      HIP_TRY(c, launch_local_dt(z, rr, c->d_tables,
                                 c->capturing ? c->d_bmin + (&b - c->blocks.data()) : c->d_scalar, s));
      break;
    }
    case COREFL_OP_WU_CN:
      if (!r.wu) return fail(c, Status::Arg, "COREFL_OP_WU_CN needs temporal_scheme 4");
      HIP_TRY(c, launch_copy(z.cn, z.dq, z.dsz * r.nv, s));
      break;
    case COREFL_OP_WU_REACTION:
      if (!r.wu) return fail(c, Status::Arg, "COREFL_OP_WU_REACTION needs temporal_scheme 4");
      if (r.reaction) {
        if (c->chem_jit.wu && r.chem_advance == 0)
          HIP_TRY(c, chem_jit_launch_wu(c->chem_jit, z, dt, nullptr, s));
        else
          HIP_TRY(c, launch_wu_reaction(z, r, c->d_tables, dt, r.chem_advance, s));
      }
      break;
    case COREFL_OP_UPDATE_RK_WU:
      if (!r.wu) return fail(c, Status::Arg, "COREFL_OP_UPDATE_RK_WU needs temporal_scheme 4");
      if (rk < 0 || rk > 2) return fail(c, Status::Arg, "rk stage must be 0, 1 or 2");
      HIP_TRY(c, launch_update_rk_wu(z, r, c->d_tables, dt, rk, s));
      break;
    case COREFL_OP_COPY_QN:
      HIP_TRY(c, launch_copy(z.qn, z.cv, z.sz * r.nv, s));
      break;
    default:
      return fail(c, Status::Arg, "unknown operator id");
  }
  return 0;
}

// Every node an inner face loop touches, on both sides, lies inside the
// padded extent of its block (the maps are affine in the loop position, so
// the corners of the loop box bound them); the target block exists.
int check_inner(corefl_ctx *c) {
  if (c->inner_checked) return 0;
  const int nb = static_cast<int>(c->blocks.size());
  for (int b = 0; b < nb; ++b) {
    const Zone &z = c->blocks[b].z;
    for (const InnerDev &f : c->blocks[b].inner) {
      if (f.target_block < 0 || f.target_block >= nb) return fail(c, Status::Arg, "inner face targets a missing block");
      const Zone &t = c->blocks[f.target_block].z;
      const int m[3] = {z.mx, z.my, z.mz}, mt[3] = {t.mx, t.my, t.mz};
      for (int a = 0; a < 3; ++a)
        if (f.n_point[a] < 1) return fail(c, Status::Arg, "empty inner face");
      for (int corner = 0; corner < 8; ++corner) {
        int idx[3], it[3], d[3];
        for (int a = 0; a < 3; ++a) {
          const int n = (corner >> a) & 1 ? f.n_point[a] - 1 : 0;
          d[a] = f.loop_dir[a] * n;
          idx[a] = f.range_start[a] + d[a];
        }
        for (int a = 0; a < 3; ++a) it[a] = f.target_start[a] + f.target_loop_dir[a] * d[f.src_tar[a]];
        for (int a = 0; a < 3; ++a) {
          const int gz = (a == 2 && z.mz == 1) ? 0 : z.ngg, gt = (a == 2 && t.mz == 1) ? 0 : t.ngg;
          if (idx[a] < -gz || idx[a] > m[a] - 1 + gz || it[a] < -gt || it[a] > mt[a] - 1 + gt)
            return fail(c, Status::Arg, "inner face range does not fit the blocks it connects");
        }
      }
    }
  }
  c->inner_checked = true;
  return 0;
}

// data_communication: inner faces of every block in block and face order,
// src/DataCommunication.cu:6-35
int halo_all(corefl_ctx *c) {
  OpTimer timer(c, COREFL_OP_HALO);
  if (int st = check_inner(c)) return st;
  for (auto &b : c->blocks)
    for (const InnerDev &f : b.inner)
      HIP_TRY(c, launch_inner_comm(b.z, c->blocks[f.target_block].z, c->run, c->d_tables, f, c->stream));
  return 0;
}

// compute_viscous_flux, 8th-order branch, src/ViscousScheme.cu:62-83: nodal
// fluxes on every block, exchange_value(task 1) over inner faces of every
// block then periodic faces of every block (src/DataCommunication.cu:437-485),
// then the derivative on every block.
int viscous_all(corefl_ctx *c) {
  OpTimer timer(c, COREFL_OP_VISCOUS);
  const Run &r = c->run;
  hipStream_t s = c->stream;
  if (r.viscous_order == 2) {
    for (auto &b : c->blocks) HIP_TRY(c, launch_viscous2(b.z, r, c->d_tables, s));
    return 0;
  }
  if (r.viscous_order != 8) return 0;
  if (int st = check_inner(c)) return st;
  for (auto &b : c->blocks) HIP_TRY(c, launch_viscous_flux(b.z, r, c->d_tables, s));
  for (auto &b : c->blocks)
    for (const InnerDev &f : b.inner) HIP_TRY(c, launch_inner_flux(b.z, c->blocks[f.target_block].z, r, f, s));
  for (auto &b : c->blocks)
    HIP_TRY(c, periodic_launches(
                   b, true, [&](const corefl_boundary &f) { return launch_periodic_flux(b.z, r, f, s); },
                   [&](const corefl_boundary &lo) { return launch_periodic_flux_pair(b.z, r, lo, s); }));
  for (auto &b : c->blocks) HIP_TRY(c, launch_viscous_derivative(b.z, r, s));
  return 0;
}

int op_all(corefl_ctx *c, int op, int rk, double dt) {
  if (op == COREFL_OP_VISCOUS) return viscous_all(c);
  if (op == COREFL_OP_HALO) return halo_all(c);
  for (auto &b : c->blocks) {
    const int st = op_block(c, op, b, rk, dt);
    if (st) return st;
  }
  return 0;
}

int global_dt(corefl_ctx *c, double *dt) {
  double m = 1e+6;
  for (auto &b : c->blocks) {
    int st = op_block(c, COREFL_OP_LOCAL_DT, b, 0, 0);
    if (st) return st;
    double v = 0;
    HIP_TRY(c, hipMemcpyAsync(&v, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    m = std::fmin(m, v);
  }
  *dt = m;
  return 0;
}

// ---- the recorded step (desc.step_graph) ----
// Every block stream waits for what `stream` has recorded so far.
int fork_blocks(corefl_ctx *c) {
  HIP_TRY(c, hipEventRecord(c->fork_ev, c->stream));
  for (hipStream_t bs : c->bstreams) HIP_TRY(c, hipStreamWaitEvent(bs, c->fork_ev, 0));
  return 0;
}

// `stream` waits for every block stream.
int join_blocks(corefl_ctx *c) {
  for (size_t b = 0; b < c->bstreams.size(); ++b) {
    HIP_TRY(c, hipEventRecord(c->join_ev[b], c->bstreams[b]));
    HIP_TRY(c, hipStreamWaitEvent(c->stream, c->join_ev[b], 0));
  }
  return 0;
}

// f(block) for every block, each block on its own stream, between a fork and a join.
template <class F>
int per_block(corefl_ctx *c, F &&f) {
  if (int st = fork_blocks(c)) return st;
  for (size_t b = 0; b < c->blocks.size(); ++b) {
    c->cur = c->bstreams[b];
    const int st = f(c->blocks[b]);
    c->cur = nullptr;
    if (st) return st;
  }
  return join_blocks(c);
}

// One stage in the order of corefl_stage with each block's work on its own
// stream.  Blocks meet only where they exchange data: after the nodal viscous
// fluxes (inner-face flux exchange) and after the boundary conditions
// (inner-face halo); without inner faces a stage is one chain per block.
int record_stage(corefl_ctx *c, int rk) {
  const Run &r = c->run;
  bool inner = false;
  for (const auto &b : c->blocks) inner = inner || !b.inner.empty();
  const bool visc8 = r.viscous_order == 8;
  int st;
  auto ops = [&](corefl_ctx::Block &b, std::initializer_list<int> list) -> int {
    for (int op : list) {
      if (op == COREFL_OP_LIMIT_FLOW && !r.limit_flow) continue;
      if ((st = op_block(c, op, b, rk, 0))) return st;
    }
    return 0;
  };
  auto viscous_head = [&](corefl_ctx::Block &b) -> int {
    hipStream_t s = c->cur;
    if (r.viscous_order == 2) HIP_TRY(c, launch_viscous2(b.z, r, c->d_tables, s));
    if (visc8) HIP_TRY(c, launch_viscous_flux(b.z, r, c->d_tables, s));
    return 0;
  };
  auto viscous_tail = [&](corefl_ctx::Block &b) -> int {
    if (!visc8) return 0;
    hipStream_t s = c->cur;
    HIP_TRY(c, periodic_launches(
                   b, true, [&](const corefl_boundary &f) { return launch_periodic_flux(b.z, r, f, s); },
                   [&](const corefl_boundary &lo) { return launch_periodic_flux_pair(b.z, r, lo, s); }));
    // the derivative is the stage's first contribution to dq and stores it (dq is not cleared first)
    HIP_TRY(c, launch_viscous_derivative(b.z, r, s, true));
    return 0;
  };
  auto rhs_update = [&](corefl_ctx::Block &b) -> int {
    if ((st = viscous_tail(b))) return st;
    return ops(b, {COREFL_OP_CHEMISTRY, COREFL_OP_INVISCID_X, COREFL_OP_INVISCID_Y, COREFL_OP_INVISCID_Z,
                   COREFL_OP_UPDATE_RK, COREFL_OP_LIMIT_FLOW, COREFL_OP_BC});
  };
  auto closing = [&](corefl_ctx::Block &b) -> int { return ops(b, {COREFL_OP_K_GRADIENT, COREFL_OP_PROPERTIES}); };
  if (inner && c->batched) {
    // the same stage with the block-local pointwise operators and the outflow faces of all blocks in one
    // launch each (blocks are independent between the flux exchange and the halo, so this is the
    // per-block order); the sweeps stay per block on the block streams
    if ((st = check_inner(c))) return st;
    const int nb = static_cast<int>(c->blocks.size());
    if (!visc8) HIP_TRY(c, launch_zero_dq_multi(c->d_zones, nb, c->mb_interior, r.nv, c->stream));
    if (r.viscous_order && (st = per_block(c, viscous_head))) return st;
    if (visc8)
      for (auto &b : c->blocks)
        for (const InnerDev &f : b.inner)
          HIP_TRY(c, launch_inner_flux(b.z, c->blocks[f.target_block].z, r, f, c->stream));
    if ((st = per_block(c, [&](corefl_ctx::Block &b) -> int {
           if ((st = viscous_tail(b))) return st;
           return ops(b, {COREFL_OP_CHEMISTRY, COREFL_OP_INVISCID_X, COREFL_OP_INVISCID_Y, COREFL_OP_INVISCID_Z});
         })))
      return st;
    HIP_TRY(c, launch_update_rk_multi(c->d_zones, nb, c->mb_interior, r, c->d_tables, c->d_scalar + 1, rk, c->stream,
                                      c->qn_in_update && rk == 0));
    for (auto &b : c->blocks)
      if ((st = ops(b, {COREFL_OP_LIMIT_FLOW}))) return st;
    // boundary conditions in their priority order: outflow (all blocks at once), then the other kinds
    HIP_TRY(c, launch_bc_outflow_multi(c->d_zones, c->d_outflow, c->n_outflow, c->mb_face, r, c->d_tables, c->stream));
    for (auto &b : c->blocks) {
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_INFLOW) HIP_TRY(c, launch_bc_inflow(b.z, r, c->d_tables, f.b, f.cond, c->stream));
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_SYMMETRY) HIP_TRY(c, launch_bc_symmetry(b.z, r, c->d_tables, f.b, c->stream));
      for (const auto &f : b.faces)
        if (f.kind == COREFL_BC_WALL) HIP_TRY(c, launch_bc_wall(b.z, r, c->d_tables, f.b, f.cond, c->stream));
      HIP_TRY(c, periodic_launches(
                     b, !r.tt,
                     [&](const corefl_boundary &f) { return launch_bc_periodic(b.z, r, c->d_tables, f, c->stream); },
                     [&](const corefl_boundary &lo) {
                       return launch_bc_periodic_pair(b.z, r, c->d_tables, lo, c->stream);
                     }));
    }
    for (auto &b : c->blocks)
      for (const InnerDev &f : b.inner)
        HIP_TRY(c, launch_inner_comm(b.z, c->blocks[f.target_block].z, r, c->d_tables, f, c->stream));
    for (auto &b : c->blocks)
      if ((st = ops(b, {COREFL_OP_K_GRADIENT}))) return st;
    HIP_TRY(c, launch_properties_multi(c->d_zones, nb, c->mb_padded, r, c->d_tables, c->stream));
    return 0;
  }
  if (!inner) {
    return per_block(c, [&](corefl_ctx::Block &b) -> int {
      if ((!visc8 && (st = ops(b, {COREFL_OP_ZERO_DQ}))) || (st = viscous_head(b)) || (st = rhs_update(b))) return st;
      return closing(b);
    });
  }
  if ((st = check_inner(c))) return st;
  if ((st = per_block(c, [&](corefl_ctx::Block &b) -> int {
         if (!visc8 && (st = ops(b, {COREFL_OP_ZERO_DQ}))) return st;
         return viscous_head(b);
       })))
    return st;
  if (visc8)
    for (auto &b : c->blocks)
      for (const InnerDev &f : b.inner)
        HIP_TRY(c, launch_inner_flux(b.z, c->blocks[f.target_block].z, r, f, c->stream));
  if ((st = per_block(c, rhs_update))) return st;
  for (auto &b : c->blocks)
    for (const InnerDev &f : b.inner)
      HIP_TRY(c, launch_inner_comm(b.z, c->blocks[f.target_block].z, r, c->d_tables, f, c->stream));
  return per_block(c, closing);
}

// Records one RK3 step (qn copy, sensor, dt, three stages) on `stream`.
int record_step(corefl_ctx *c, double dt, bool fixed) {
  int st;
  // the first stage's update saves qn from cv (the step's start state) instead of a copy of cv here
  struct QnInUpdate {
    corefl_ctx *c;
    explicit QnInUpdate(corefl_ctx *ctx) : c(ctx) { c->qn_in_update = true; }
    ~QnInUpdate() { c->qn_in_update = false; }
  } qn_guard(c);
  if ((st = per_block(c, [&](corefl_ctx::Block &b) -> int {
         if (c->run.hybrid && (st = op_block(c, COREFL_OP_SHOCK_SENSOR, b, 0, 0))) return st;
         if (!fixed && (st = op_block(c, COREFL_OP_LOCAL_DT, b, 0, 0))) return st;
         return 0;
       })))
    return st;
  if (fixed)
    HIP_TRY(c, launch_fill(c->d_scalar + 1, 1, dt, c->stream));
  else
    HIP_TRY(c, launch_min_combine(c->d_bmin, static_cast<int>(c->blocks.size()), c->d_scalar + 1, c->stream));
  for (int rk = 0; rk < 3; ++rk)
    if ((st = record_stage(c, rk))) return st;
  return 0;
}

void drop_graph(corefl_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  c->graph_exec = nullptr;
  if (c->d_zones) (void)hipFree(c->d_zones);
  if (c->d_outflow) (void)hipFree(c->d_outflow);
  c->d_zones = nullptr;
  c->d_outflow = nullptr;
  c->batched = false;
}

// Device tables for the multi-block launches (one-temperature runs with several blocks, none above
// kBatchPoints interior points: larger blocks fill the GPU alone and gain more from running their
// chains side by side on the block streams).
constexpr long kBatchPoints = 1L << 20;

int upload_block_tables(corefl_ctx *c) {
  const size_t nb = c->blocks.size();
  if (nb < 2 || c->run.tt) return 0;
  for (const auto &b : c->blocks)
    if (static_cast<long>(b.z.mx) * b.z.my * b.z.mz > kBatchPoints) return 0;
  std::vector<Zone> zs;
  std::vector<BcItem> out;
  c->mb_interior = c->mb_padded = c->mb_face = 0;
  for (size_t i = 0; i < nb; ++i) {
    const Zone &z = c->blocks[i].z;
    zs.push_back(z);
    c->mb_interior = std::max<long>(c->mb_interior, static_cast<long>(z.mx) * z.my * z.mz);
    c->mb_padded = std::max<long>(c->mb_padded, static_cast<long>(z.mx + 2 * z.ngg) * (z.my + 2 * z.ngg) *
                                                    (z.mz + 2 * z.ngg));
    for (const auto &f : c->blocks[i].faces) {
      if (f.kind != COREFL_BC_OUTFLOW) continue;
      BcItem it;
      it.blk = static_cast<int>(i);
      it.b = f.b;
      out.push_back(it);
      long n = 1;
      for (int d = 0; d < 3; ++d) n *= f.b.range_end[d] - f.b.range_start[d] + 1;
      c->mb_face = std::max(c->mb_face, n);
    }
  }
  HIP_TRY(c, hipMalloc(&c->d_zones, nb * sizeof(Zone)));
  HIP_TRY(c, hipMemcpy(c->d_zones, zs.data(), nb * sizeof(Zone), hipMemcpyHostToDevice));
  c->n_outflow = static_cast<int>(out.size());
  HIP_TRY(c, hipMalloc(&c->d_outflow, std::max<size_t>(out.size(), 1) * sizeof(BcItem)));
  if (!out.empty())
    HIP_TRY(c, hipMemcpy(c->d_outflow, out.data(), out.size() * sizeof(BcItem), hipMemcpyHostToDevice));
  c->batched = true;
  return 0;
}

// Builds (or rebuilds) the step graph: block streams and events first, then
// the capture of record_step on `stream`.
int build_graph(corefl_ctx *c, double dt, bool fixed) {
  drop_graph(c);
  const size_t nb = c->blocks.size();
  while (c->bstreams.size() < nb) {
    hipStream_t s;
    HIP_TRY(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    c->bstreams.push_back(s);
    hipEvent_t e;
    HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->join_ev.push_back(e);
  }
  if (!c->fork_ev) HIP_TRY(c, hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
  if (!c->d_bmin || c->graph_nb < nb) {
    if (c->d_bmin) HIP_TRY(c, hipFree(c->d_bmin));
    c->d_bmin = nullptr;
    HIP_TRY(c, hipMalloc(&c->d_bmin, std::max<size_t>(nb, 1) * sizeof(double)));
    c->graph_nb = nb;
  }
  if (int st0 = upload_block_tables(c)) return st0;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  const int st = record_step(c, dt, fixed);
  c->capturing = false;
  c->cur = nullptr;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  HIP_TRY(c, e);
  const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIP_TRY(c, ei);
  c->graph_dt = dt;
  return 0;
}

// The update kernels' state flag (src/Residual.cu:97-118 stops the run on such a state): a copy to pinned
// host memory queued behind the work just issued, with an event, so a later call reads it without waiting.
int post_state_flag(corefl_ctx *c) {
  HIP_TRY(c, hipMemcpyAsync(c->h_bad, c->d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipEventRecord(c->bad_ev, c->stream));
  c->bad_posted = true;
  return 0;
}

// wait: copy the flag now and wait for it; otherwise read the last posted copy if it has landed.
int check_state_flag(corefl_ctx *c, bool wait) {
  if (wait) {
    if (int st = post_state_flag(c)) return st;
    HIP_TRY(c, hipEventSynchronize(c->bad_ev));
  } else {
    if (!c->bad_posted) return 0;
    const hipError_t q = hipEventQuery(c->bad_ev);
    if (q == hipErrorNotReady) {
      (void)hipGetLastError();  // not-ready is an answer, not an error to keep
      return 0;
    }
    HIP_TRY(c, q);
  }
  if (*c->h_bad)
    return fail(c, Status::State,
                "the RK update produced a non-finite or non-positive density, pressure or temperature, or a "
                "non-finite velocity or species mass fraction");
  return 0;
}

// Clears the flag before a new state (corefl_state_upload, corefl_initialize).
int clear_state_flag(corefl_ctx *c) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemsetAsync(c->d_bad, 0, sizeof(int), c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *c->h_bad = 0;
  c->bad_posted = false;
  return 0;
}

int run_stage(corefl_ctx *c, int rk, double dt);
int run_step(corefl_ctx *c, double *dt_inout);

}  // namespace

void derive_tables(const corefl_chem_tables &in, Tab &out) {
  static_cast<corefl_chem_tables &>(out) = in;
  for (int l = 0; l < in.n_spec; ++l) {
    const double ln_lj = in.LJ_potent_inv[l] > 0 ? std::log(in.LJ_potent_inv[l]) : 0.0;
    out.pw_lj[l] = std::exp(-0.145 * ln_lj);
    for (int k = 0; k < in.n_spec; ++k) {
      const double ln_kbe = in.kb_over_eps_jk[l][k] > 0 ? std::log(in.kb_over_eps_jk[l][k]) : 0.0;
      out.pw_kbe[l][k] = 1.0548 * std::exp(-0.15504 * ln_kbe);
      out.inv_bdc[l][k] = in.binary_diffusivity_coeff[l][k] != 0 ? 1.0 / in.binary_diffusivity_coeff[l][k] : 0.0;
    }
  }
  for (int r = 0; r < in.n_reac; ++r) {
    Reac &R = out.reac[r];
    R = Reac{};
    for (int l = 0; l < in.n_spec; ++l) {
      const int sf = in.stoi_f[r][l], sb = in.stoi_b[r][l];
      if (sf != 0) {
        R.f_idx[R.nf] = l;
        R.f_ord[R.nf++] = sf;
      }
      if (sb != 0) {
        R.b_idx[R.nb] = l;
        R.b_ord[R.nb++] = sb;
      }
      if (sb - sf != 0) {
        R.nu_idx[R.nnu] = l;
        R.nu[R.nnu++] = static_cast<double>(sb - sf);
      }
      if (sf != 0 || sb != 0) {
        R.part_idx[R.npart] = l;
        R.part_f[R.npart] = static_cast<double>(sf);
        R.part_b[R.npart++] = static_cast<double>(sb);
      }
    }
  }
}

static int sgn_of(int a) { return a < 0 ? -1 : 1; }  // gxl::sgn, src/gxl_lib/Math.hpp:10-12

// InnerFace::register_boundary, src/Mesh.cpp:39-172: the face axis is the one
// whose range collapses (value 1: the low face), the negated tangential axes
// of the two lines correspond to each other, and the order of start and end
// gives the running direction.  The loop along the normal covers the face
// node and the ghost layers.
void register_inner(const corefl_inner_face &in, int dim, int ngg, InnerDev &f) {
  int rs[3] = {in.src[0], in.src[2], in.src[4]}, re[3] = {in.src[1], in.src[3], in.src[5]};
  int ts[3] = {in.tar[0], in.tar[2], in.tar[4]}, te[3] = {in.tar[1], in.tar[3], in.tar[5]};
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
    rs[nf] = std::abs(rs[nf]) - 1;
    re[nf] = std::abs(re[nf]) - 1;
    ts[tnf] = std::abs(ts[tnf]) - 1;
    te[tnf] = std::abs(te[tnf]) - 1;
    st[tnf] = sgn_of((te[tnf] - ts[tnf]) * (re[nf] - rs[nf])) * (nf + 1);
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
    st[tnf] = sgn_of((te[tnf] - ts[tnf]) * (re[nf] - rs[nf])) * (nf + 1);
    nf = 3 - face - nf;
    tnf = 3 - tface - tnf;
    --rs[nf];
    --re[nf];
    --ts[tnf];
    --te[tnf];
    st[tnf] = sgn_of((te[tnf] - ts[tnf]) * (re[nf] - rs[nf])) * (nf + 1);
  }
  rs[face] = dirn < 0 ? -ngg : re[face];
  re[face] = dirn < 0 ? 0 : re[face] + ngg;
  const int sgn_tar = sgn_of(st[tface]);
  if (tdirn > 0)
    ts[tface] = sgn_tar > 0 ? te[tface] - ngg : te[tface];
  else
    ts[tface] = sgn_tar > 0 ? 0 : ngg;
  for (int a = 0; a < 3; ++a) {
    f.range_start[a] = rs[a];
    f.range_end[a] = re[a];
    f.target_start[a] = ts[a];
    f.loop_dir[a] = rs[a] <= re[a] ? 1 : -1;
    f.target_loop_dir[a] = sgn_of(st[a]);
    f.src_tar[a] = std::abs(st[a]) - 1;
    f.n_point[a] = std::abs(rs[a] - re[a]) + 1;
  }
  f.face = face;
  f.direction = dirn;
  f.target_face = tface;
  f.target_direction = tdirn;
  f.target_block = in.target_block - 1;
}

}  // namespace corefl

using namespace corefl;

extern "C" {

int corefl_chem_load(const char *mech, const char *therm, const char *tran, corefl_chem_tables *out) {
  if (!mech || !therm || !tran || !out) return static_cast<int>(Status::Arg);
  std::string err;
  const Status s = load_chemistry(mech, therm, tran, *out, err);
  if (s != Status::Ok) std::fprintf(stderr, "corefl_chem_load: %s\n", err.c_str());
  return static_cast<int>(s);
}

int corefl_chem_load_ex(const char *mech, const char *therm, const char *tran, int nasa_7_or_9,
                        const char *two_temperature, corefl_chem_tables *out) {
  if (!mech || !therm || !tran || !out) return static_cast<int>(Status::Arg);
  std::string err;
  const Status s = load_chemistry_ex(mech, therm, tran, nasa_7_or_9, two_temperature ? two_temperature : "", *out, err);
  if (s != Status::Ok) std::fprintf(stderr, "corefl_chem_load_ex: %s\n", err.c_str());
  return static_cast<int>(s);
}

int corefl_ctx_create(const corefl_run_desc *run, const corefl_chem_tables *tables, corefl_ctx **out) {
  if (!run || !out) return static_cast<int>(Status::Arg);
  *out = nullptr;
  auto *c = new corefl_ctx();
  c->desc = *run;
  Run &r = c->run;
  r.mixture = run->mixture;
  r.ns = run->mixture ? run->n_spec : 0;
  r.model = r.ns;
  r.nv = run->n_var;
  r.n_scalar = run->n_scalar;
  r.nst = run->n_scalar_transported;
  r.scheme = run->inviscid_scheme;
  r.viscous_order = run->viscous_order;
  r.reaction = run->reaction;
  r.gradp = run->gradp_in_diffusion;
  r.prandtl = run->prandtl;
  r.cfl = run->cfl;
  // eps_ref = 1e-40 * (rho_ref v_ref)^2, src/WENO.cu:884 and src/DParameter.cu:29-31
  r.eps_ref = 1e-40 * (run->rho_ref * run->v_ref * run->rho_ref * run->v_ref);
  auto bad = [&](const char *m) {
    c->err = m;
    *out = c;
    return static_cast<int>(Status::Arg);
  };
  if (r.mixture) {
    if (!tables) return bad("mixture runs need chemistry tables");
    if (tables->n_spec != r.ns) return bad("n_spec differs from the tables");
    if (tables->n_reac < 0 || tables->n_reac > COREFL_MAX_REAC || tables->n_spec > COREFL_MAX_SPEC)
      return bad("table sizes exceed COREFL_MAX_SPEC / COREFL_MAX_REAC");
    if (tables->nasa7or9 != 7 && tables->nasa7or9 != 9) return bad("nasa7or9 must be 7 or 9");
    r.model = r.ns | (tables->nasa7or9 == 9 ? 32 : 0);
    if (!model_supported(r.model))
      return bad("no kernels compiled for this mixture (supported: 5 or 9 species with NASA-7, 5 with NASA-9)");
    r.tt = run->two_temperature ? 1 : 0;
    r.vt = run->vt_relaxation ? 1 : 0;
    if (r.tt != (tables->two_temperature ? 1 : 0))
      return bad("two_temperature runs need tables loaded with two-temperature data, and only they");
    if (r.tt) {
      r.model |= 64;
      if (!model_supported(r.model))
        return bad("the two-temperature kernels are compiled for 5 species with NASA-9 thermodynamics");
      if (r.nv != 6 + r.ns || r.n_scalar != r.ns + 1 || r.nst != r.ns + 1)
        return bad("two-temperature runs have n_var 6+n_spec and n_scalar n_spec+1 (E_ve last)");
    } else if (r.nv != 5 + r.ns || r.n_scalar != r.ns || r.nst != r.ns) {
      return bad("n_var/n_scalar must be 5+n_spec/n_spec for the one-temperature mixture path");
    }
  } else {
    if (run->two_temperature) return bad("the two-temperature model needs a species mixture");
    if (r.nv != 5 || r.n_scalar != 0) return bad("Air runs have n_var 5 and no scalars");
  }
  r.recon = run->reconstruction;
  r.limiter = run->limiter;
  r.efix = run->entropy_fix_factor;
  // limit_flow bounds, src/DParameter.cu:371-387 (v_ref is already the sound speed where |v| < 1)
  r.limit_flow = run->limit_flow;
  r.ll[0] = 1e-6 * run->rho_ref;
  r.ul[0] = 1e+3 * run->rho_ref;
  for (int l = 1; l < 4; ++l) {
    r.ll[l] = -1000 * run->v_ref;
    r.ul[l] = 1000 * run->v_ref;
  }
  r.ll[4] = 1e-6 * run->p_ref;
  r.ul[4] = 1e+3 * run->p_ref;
  const bool weno = r.scheme == 51 || r.scheme == 52 || r.scheme == 71 || r.scheme == 72;
  if (!weno && (r.scheme < 1 || r.scheme > 4))
    return bad("inviscid_scheme must be 52 or 72 (characteristic WENO), 51 or 71 (component-wise WENO), 1 (LF), "
               "2 (Roe), 3 (AUSM+) or 4 (HLLC)");
  // WENO options, src/DParameter.cu:13, 52-54 and src/Parameter.cpp:252-270
  r.pp = run->positive_preserving ? 1 : 0;
  r.pp_dt = run->fixed_time_step ? run->dt : 0.0;
  r.hybrid = weno && run->sensor_threshold > 1e-10;
  r.sensor_type = run->shock_sensor;
  r.sensor_eps = run->sensor_eps;
  r.sensor_threshold = run->sensor_threshold;
  r.rho_ref = run->rho_ref;
  r.v_ref = run->v_ref;
  if (r.hybrid && (r.sensor_type < 0 || r.sensor_type > 2))
    return bad("shock_sensor must be 0 (Ducros), 1 (Jameson) or 2 (density/pressure jump)");
  if (r.scheme == 1 && r.mixture) return bad("the Lax-Friedrichs flux is for the Air perfect gas only");
  if (r.scheme <= 4) {
    if (r.recon < 1 || r.recon > 5)
      return bad("reconstruction must be 1 (first order), 2 (MUSCL), 3 (NND2), 4 (WENO5) or 5 (WENO7)");
    if (r.limiter != 0 && r.limiter != 1) return bad("limiter must be 0 or 1 (minmod)");
    if (r.scheme == 2 && !(r.efix > 0)) return bad("the Roe flux needs a positive entropy_fix_factor");
  }
  if (r.viscous_order != 0 && r.viscous_order != 2 && r.viscous_order != 8)
    return bad("viscous_order must be 0, 2 or 8");
  if (r.reaction && !r.mixture) return bad("reaction needs a species mixture");
  if (run->chem_kernel != 0 && run->chem_kernel != 1)
    return bad("chem_kernel must be 0 (compiled for the mechanism) or 1 (table-driven)");
  if (!run->fixed_time_step && !(run->cfl > 0)) return bad("CFL-based time step needs cfl > 0");
  const int ts = run->temporal_scheme == 0 ? 3 : run->temporal_scheme;
  if (ts != 3 && ts != 4) return bad("temporal_scheme must be 3 (SSP-RK3) or 4 (Wu's operator splitting)");
  r.wu = ts == 4 ? 1 : 0;
  r.chem_advance = run->chem_advance_method;
  if (r.chem_advance != 0 && r.chem_advance != 1 && r.chem_advance != 3)
    return bad("chem_advance_method must be 0 (explicit), 1 (point-implicit) or 3 (SSP-RK3)");
  if (r.wu && (!r.mixture || r.tt || !r.reaction))
    return bad("Wu's operator splitting runs a one-temperature reacting mixture");
  if (r.wu && run->step_graph) return bad("the recorded step covers SSP-RK3; Wu's splitting runs operator by operator");
  if (tables) c->host_tables = *tables;
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->d_tables, sizeof(Tab));
  auto dev_tables = std::make_unique<Tab>();
  derive_tables(c->host_tables, *dev_tables);
  if (e == hipSuccess) e = hipMemcpy(c->d_tables, dev_tables.get(), sizeof(Tab), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->d_scalar, 16 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c->d_bad, sizeof(int));
  if (e == hipSuccess) e = hipMemset(c->d_bad, 0, sizeof(int));
  if (e == hipSuccess) e = hipHostMalloc(&c->h_bad, sizeof(int), hipHostMallocDefault);
  if (e == hipSuccess) *c->h_bad = 0;
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->bad_ev, hipEventDisableTiming);
  *out = c;
  if (e != hipSuccess) {
    c->err = std::string("HIP setup failed: ") + hipGetErrorString(e);
    return static_cast<int>(Status::Hip);
  }
  // the compiled-per-mechanism chemistry covers NASA-7 one-temperature mixtures; others use the table kernel
  if (r.mixture && r.reaction && run->chem_kernel == 0 && r.model == r.ns) {
    std::string err;
    const Status st = chem_jit_build(*dev_tables, r.ns, c->chem_jit, err);
    if (st != Status::Ok) return fail(c, st, "compiling the chemistry kernel for the mechanism: " + err);
  }
  return 0;
}

void corefl_ctx_destroy(corefl_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (hipStream_t s : c->bstreams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : c->join_ev) (void)hipEventDestroy(e);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->d_bmin) (void)hipFree(c->d_bmin);
  for (auto &b : c->blocks)
    for (void *p : b.allocs) (void)hipFree(p);
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  if (c->d_tables) (void)hipFree(c->d_tables);
  if (c->d_scalar) (void)hipFree(c->d_scalar);
  if (c->d_bad) (void)hipFree(c->d_bad);
  if (c->h_bad) (void)hipHostFree(c->h_bad);
  if (c->bad_ev) (void)hipEventDestroy(c->bad_ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *corefl_last_error(const corefl_ctx *c) { return c ? c->err.c_str() : "null context"; }

int corefl_bc_set(corefl_ctx *c, const corefl_bc_condition *cond) {
  if (!c) return static_cast<int>(Status::Arg);
  if (!cond) return fail(c, Status::Arg, "null boundary condition");
  switch (cond->kind) {
    case COREFL_BC_WALL:
      if (cond->thermal_type != 0 && cond->thermal_type != 1)
        return fail(c, Status::Arg, "wall thermal_type must be 0 (isothermal) or 1 (adiabatic)");
      if (cond->thermal_type == 0 && !(cond->wall_temperature > 0))
        return fail(c, Status::Arg, "isothermal wall needs a positive temperature");
      break;
    case COREFL_BC_INFLOW:
      if (!(cond->density > 0) || !(cond->pressure > 0) || !(cond->temperature > 0))
        return fail(c, Status::Arg, "inflow needs positive density, pressure and temperature");
      if (c->run.mixture && cond->inflow_sub && !(cond->mw > 0))
        return fail(c, Status::Arg, "subsonic mixture inflow needs the inflow molecular weight");
      break;
    case COREFL_BC_SYMMETRY:
    case COREFL_BC_OUTFLOW:
    case COREFL_BC_PERIODIC:
      break;
    default:
      return fail(c, Status::Arg, "boundary kind must be 2 (wall), 3 (symmetry), 5 (inflow), 6 (outflow) or 8 (periodic)");
  }
  corefl_bc_condition cc = *cond;
  if (c->run.tt && cc.kind == COREFL_BC_INFLOW) {
    // Inflow's constructor with two temperatures (src/BoundCond.cpp:256-262): E_ve of the stream at its Tve
    const double tve = cc.temperature_ve > 0 ? cc.temperature_ve : cc.temperature;
    double eve = 0;
    for (int l = 0; l < c->run.ns; ++l) eve += cc.sv[l] * host_ve_energy(c->host_tables, l, tve);
    cc.sv[c->run.ns] = eve;
  }
  for (auto &cd : c->conditions)
    if (cd.label == cc.label) {
      cd = cc;
      return 0;
    }
  c->conditions.push_back(cc);
  return 0;
}

int corefl_block_add(corefl_ctx *c, const corefl_block_desc *d, int *block_id) {
  if (!c) return static_cast<int>(Status::Arg);
  if (!d || !d->metric || !d->jac) return fail(c, Status::Arg, "block descriptor needs metric and jac");
  const Run &r = c->run;
  if (d->mx < 1 || d->my < 1 || d->mz < 1 || d->ngg < 1) return fail(c, Status::Arg, "bad block extents");
  const bool pv = r.scheme >= 1 && r.scheme <= 4;
  const bool w7 = r.scheme == 72 || r.scheme == 71 || (pv && r.recon == 5),
             w5 = r.scheme == 52 || r.scheme == 51 || (pv && r.recon == 4);
  const int need = w7 || r.viscous_order == 8 ? 4 : (w5 ? 3 : 2);
  if (d->ngg < need) return fail(c, Status::Arg, "ngg too small for the schemes (WENO7 / 8th-order viscous need 4)");
  // The sweeps read at most ngg nodes beyond either end of a line, inside the padding, whatever the
  // extent; the mirrored boundary conditions take ghost layer g from node g, so a line holds ngg + 1 nodes.
  const int least = d->ngg + 1;
  if (d->mx < least || d->my < least || (d->mz > 1 && d->mz < least))
    return fail(c, Status::Arg, "block extent below ngg + 1 nodes along a direction");
  corefl_ctx::Block b;
  Zone &z = b.z;
  z.bad = c->d_bad;
  z.mx = d->mx;
  z.my = d->my;
  z.mz = d->mz;
  z.ngg = d->ngg;
  const long g = d->ngg;
  z.sx = d->mx + 2 * g;
  z.sxy = z.sx * (d->my + 2 * g);
  z.sz = z.sxy * (d->mz + 2 * g);
  z.off = g * (z.sxy + z.sx + 1);
  z.dsz = static_cast<long>(d->mx) * d->my * d->mz;
  for (int f = 0; f < d->n_boundary; ++f) {
    const corefl_boundary &bf = d->boundary[f];
    corefl_ctx::Face face{bf, 0, corefl_bc_condition{}};
    bool found = false;
    for (const auto &cd : c->conditions)
      if (cd.label == bf.type) {
        face.cond = cd;
        face.kind = cd.kind;
        found = true;
      }
    if (!found) {
      face.cond.label = bf.type;
      switch (bf.type) {
        case COREFL_BC_WALL:  // Wall defaults: isothermal at 300 K (src/BoundCond.h:102-110)
          face.kind = COREFL_BC_WALL;
          face.cond.kind = COREFL_BC_WALL;
          face.cond.thermal_type = 0;
          face.cond.wall_temperature = 300;
          break;
        case COREFL_BC_SYMMETRY:
        case COREFL_BC_OUTFLOW:
        case COREFL_BC_PERIODIC:
          face.kind = bf.type;
          face.cond.kind = bf.type;
          break;
        default:
          return fail(c, Status::Arg, "boundary label " + std::to_string(bf.type) +
                                          " has no condition (register one with corefl_bc_set)");
      }
    }
    if (bf.face < 0 || bf.face > 2 || (bf.direction != 1 && bf.direction != -1))
      return fail(c, Status::Arg, "bad boundary face/direction");
    const int n[3] = {d->mx, d->my, d->mz};
    for (int a = 0; a < 3; ++a)
      if (bf.range_start[a] < 0 || bf.range_end[a] >= n[a] || bf.range_end[a] < bf.range_start[a])
        return fail(c, Status::Arg, "boundary range outside the block interior");
    const int at = bf.direction < 0 ? 0 : n[bf.face] - 1;
    if (bf.range_start[bf.face] != at || bf.range_end[bf.face] != at)
      return fail(c, Status::Arg, "boundary face does not lie on the block face");
    b.faces.push_back(face);
  }
  if (d->n_inner < 0 || (d->n_inner > 0 && !d->inner)) return fail(c, Status::Arg, "bad inner face list");
  for (int f = 0; f < d->n_inner; ++f) {
    InnerDev id{};
    register_inner(d->inner[f], d->mz == 1 ? 2 : 3, d->ngg, id);
    b.inner.push_back(id);
  }
  c->inner_checked = false;
  drop_graph(c);
  auto alloc = [&](double *&p, long n) -> hipError_t {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(double));
    if (e != hipSuccess) return e;
    b.allocs.push_back(q);
    p = static_cast<double *>(q);
    return hipMemset(q, 0, n * sizeof(double));
  };
  const long sz = z.sz, dsz = z.dsz;
  const long nsc = r.n_scalar > 0 ? r.n_scalar : 1, nsr = r.ns > 0 ? r.ns : 1;
  const long nblk = (dsz + 255) / 256;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = alloc(z.cv, sz * r.nv);
  if (e == hipSuccess) e = alloc(z.qn, sz * r.nv);
  if (e == hipSuccess) e = alloc(z.bv, sz * 6);
  if (e == hipSuccess) e = alloc(z.sv, sz * nsc);
  if (e == hipSuccess) e = alloc(z.dq, dsz * r.nv);
  if (e == hipSuccess) e = alloc(z.ff, sz * r.nv);
  if (e == hipSuccess) e = alloc(z.gf, sz * r.nv);
  if (e == hipSuccess) e = alloc(z.hf, sz * r.nv);
  if (e == hipSuccess) e = alloc(z.metric, sz * 9);
  if (e == hipSuccess) e = alloc(z.jac, sz);
  if (e == hipSuccess) e = alloc(z.mul, sz);
  if (e == hipSuccess) e = alloc(z.kappa, sz);
  if (e == hipSuccess) e = alloc(z.gam, sz);
  if (e == hipSuccess) e = alloc(z.acoustic, sz);
  if (e == hipSuccess) e = alloc(z.mach, sz);
  if (e == hipSuccess) e = alloc(z.rhoD, sz * nsr);
  if (e == hipSuccess) e = alloc(z.rts, sz);
  if (e == hipSuccess) e = alloc(z.dt_local, dsz + nblk);
  if (e == hipSuccess && r.viscous_order == 8)
    e = alloc(z.zg, dsz * (4 + (r.ns > 0 ? r.ns + 1 : 0) + (r.tt ? 1 : 0)));
  if (e == hipSuccess && r.tt) e = alloc(z.tve, sz);
  if (e == hipSuccess && r.tt) e = alloc(z.kappa_ve, sz);
  if (e == hipSuccess && r.hybrid) e = alloc(z.sensor, sz);
  if (e == hipSuccess && r.wu) e = alloc(z.cn, dsz * r.nv);
  // bType masks, src/Mesh.cpp:856-916: physical faces carry their label, periodic faces stay 0
  std::vector<int> il(static_cast<size_t>(d->my) * d->mz, 0), ir(il), jl(static_cast<size_t>(d->mx) * d->mz, 0), jr(jl),
      kl(static_cast<size_t>(d->mx) * d->my, 0), kr(kl);
  for (const auto &fc : b.faces) {
    const corefl_boundary &bf = fc.b;
    if (fc.kind == COREFL_BC_PERIODIC) continue;
    for (int k = bf.range_start[2]; k <= bf.range_end[2]; ++k)
      for (int j = bf.range_start[1]; j <= bf.range_end[1]; ++j)
        for (int i = bf.range_start[0]; i <= bf.range_end[0]; ++i) {
          if (bf.face == 0) (bf.direction < 0 ? il : ir)[j + static_cast<size_t>(k) * d->my] = bf.type;
          if (bf.face == 1) (bf.direction < 0 ? jl : jr)[i + static_cast<size_t>(k) * d->mx] = bf.type;
          if (bf.face == 2) (bf.direction < 0 ? kl : kr)[i + static_cast<size_t>(j) * d->mx] = bf.type;
        }
  }
  auto ialloc = [&](int *&p, const std::vector<int> &h) -> hipError_t {
    void *q = nullptr;
    hipError_t ee = hipMalloc(&q, h.size() * sizeof(int));
    if (ee != hipSuccess) return ee;
    b.allocs.push_back(q);
    p = static_cast<int *>(q);
    return hipMemcpy(q, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice);
  };
  if (e == hipSuccess) e = ialloc(z.bt_il, il);
  if (e == hipSuccess) e = ialloc(z.bt_ir, ir);
  if (e == hipSuccess) e = ialloc(z.bt_jl, jl);
  if (e == hipSuccess) e = ialloc(z.bt_jr, jr);
  if (e == hipSuccess) e = ialloc(z.bt_kl, kl);
  if (e == hipSuccess) e = ialloc(z.bt_kr, kr);
  if (e == hipSuccess) e = hipMemcpy(z.metric, d->metric, sz * 9 * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(z.jac, d->jac, sz * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    for (void *p : b.allocs) (void)hipFree(p);
    return hip_fail(c, e, "block allocation");
  }
  c->blocks.push_back(std::move(b));
  if (block_id) *block_id = static_cast<int>(c->blocks.size()) - 1;
  return 0;
}

int corefl_state_upload(corefl_ctx *c, int blk, const double *bv, const double *sv) {
  if (!c) return static_cast<int>(Status::Arg);
  if (blk < 0 || blk >= static_cast<int>(c->blocks.size())) return fail(c, Status::Arg, "bad block id");
  if (!bv || (c->run.n_scalar > 0 && !sv)) return fail(c, Status::Arg, "null state array");
  const Zone &z = c->blocks[blk].z;
  HIP_TRY(c, hipMemcpyAsync(z.bv, bv, z.sz * 6 * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->run.n_scalar > 0)
    HIP_TRY(c, hipMemcpyAsync(z.sv, sv, z.sz * c->run.n_scalar * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return clear_state_flag(c);
}

long corefl_field_size(const corefl_ctx *c, int blk, int f) {
  if (!c || blk < 0 || blk >= static_cast<int>(c->blocks.size()) || f < 0 || f >= COREFL_F_COUNT) return -1;
  return field_doubles(c, c->blocks[blk].z, f);
}

int corefl_field_download(corefl_ctx *c, int blk, int f, double *out) {
  if (!c) return static_cast<int>(Status::Arg);
  if (blk < 0 || blk >= static_cast<int>(c->blocks.size()) || f < 0 || f >= COREFL_F_COUNT || !out)
    return fail(c, Status::Arg, "bad field download arguments");
  const Zone &z = c->blocks[blk].z;
  if (!field_ptr(z, f)) return fail(c, Status::Arg, "this field is not allocated for this run");
  HIP_TRY(c, hipMemcpyAsync(out, field_ptr(z, f), field_doubles(c, z, f) * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return 0;
}

int corefl_field_upload(corefl_ctx *c, int blk, int f, const double *in) {
  if (!c) return static_cast<int>(Status::Arg);
  if (blk < 0 || blk >= static_cast<int>(c->blocks.size()) || f < 0 || f >= COREFL_F_COUNT || !in)
    return fail(c, Status::Arg, "bad field upload arguments");
  const Zone &z = c->blocks[blk].z;
  if (!field_ptr(z, f)) return fail(c, Status::Arg, "this field is not allocated for this run");
  HIP_TRY(c, hipMemcpyAsync(field_ptr(z, f), in, field_doubles(c, z, f) * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return 0;
}

namespace {

// Autotuner of the characteristic WENO sweeps: each tile shape runs on the
// block's own extents (one warm-up, then kTuneRuns timed launches on the
// context's stream) and the fastest is kept per block and direction.  The
// launches only add into dq, which the tuner clears afterwards (every stage
// clears it again first), and every shape computes the same fluxes.
constexpr int kTuneRuns = 3;

int tune_blocks(corefl_ctx *c) {
  const Run &r = c->run;
  if (r.scheme != 52 && r.scheme != 72) return 0;
  hipEvent_t a = nullptr, e = nullptr;
  HIP_TRY(c, hipEventCreate(&a));
  HIP_TRY(c, hipEventCreate(&e));
  int st = 0;
  for (auto &b : c->blocks) {
    for (int dir = 0; dir < (b.z.mz == 1 ? 2 : 3) && !st; ++dir) {
      float best = 0.0f;
      for (int shape = 0; shape < kWenoShapes; ++shape) {
        hipError_t h = launch_inviscid(b.z, r, c->d_tables, dir, c->stream, shape);
        if (h == hipSuccess) h = hipEventRecord(a, c->stream);
        for (int k = 0; k < kTuneRuns && h == hipSuccess; ++k)
          h = launch_inviscid(b.z, r, c->d_tables, dir, c->stream, shape);
        if (h == hipSuccess) h = hipEventRecord(e, c->stream);
        if (h == hipSuccess) h = hipEventSynchronize(e);
        float ms = 0.0f;
        if (h == hipSuccess) h = hipEventElapsedTime(&ms, a, e);
        if (h != hipSuccess) {
          st = fail(c, Status::Hip, std::string("tuning the WENO tile shapes: ") + hipGetErrorString(h));
          break;
        }
        if (shape == 0 || ms < best) {
          best = ms;
          b.shape[dir] = shape;
        }
      }
    }
    if (!st) {
      const hipError_t h = launch_fill(b.z.dq, b.z.dsz * r.nv, 0.0, c->stream);
      if (h != hipSuccess) st = fail(c, Status::Hip, hipGetErrorString(h));
    }
    if (st) break;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(e);
  return st;
}

}  // namespace

int corefl_initialize(corefl_ctx *c) {
  if (!c) return static_cast<int>(Status::Arg);
  int st;
  if ((st = clear_state_flag(c))) return st;
  if ((st = op_all(c, COREFL_OP_K_GRADIENT, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_PROPERTIES, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_BC, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_HALO, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_PROPERTIES, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_CV_FROM_BV, 0, 0))) return st;
  if (c->desc.autotune && (st = tune_blocks(c))) return st;
  return 0;
}

int corefl_block_set_tiles(corefl_ctx *c, int blk, const int shapes[3]) {
  if (!c || !shapes || blk < 0 || blk >= static_cast<int>(c->blocks.size())) return static_cast<int>(Status::Arg);
  for (int d = 0; d < 3; ++d)
    if (shapes[d] < 0 || shapes[d] >= kWenoShapes) return fail(c, Status::Arg, "tile shapes are 0 or 1");
  for (int d = 0; d < 3; ++d) c->blocks[blk].shape[d] = shapes[d];
  drop_graph(c);  // a recorded step holds the old launches
  return 0;
}

int corefl_block_tiles(const corefl_ctx *c, int blk, int shapes[3]) {
  if (!c || !shapes || blk < 0 || blk >= static_cast<int>(c->blocks.size())) return static_cast<int>(Status::Arg);
  for (int d = 0; d < 3; ++d) shapes[d] = c->blocks[blk].shape[d];
  return 0;
}

int corefl_op(corefl_ctx *c, int op, int blk, int rk, double dt) {
  if (!c) return static_cast<int>(Status::Arg);
  if (op == COREFL_OP_LOCAL_DT) {
    double v;
    return global_dt(c, &v);
  }
  if (blk < 0 || op == COREFL_OP_VISCOUS || op == COREFL_OP_HALO) return op_all(c, op, rk, dt);
  if (blk >= static_cast<int>(c->blocks.size())) return fail(c, Status::Arg, "bad block id");
  return op_block(c, op, c->blocks[blk], rk, dt);
}

int corefl_stage(corefl_ctx *c, int rk, double dt) {
  if (!c) return static_cast<int>(Status::Arg);
  int st;
  if (c->run.wu) return fail(c, Status::Arg, "corefl_stage is the SSP-RK3 stage; Wu's splitting advances by corefl_step");
  if ((st = check_state_flag(c, false))) return st;
  if ((st = run_stage(c, rk, dt))) return st;
  return post_state_flag(c);
}

int corefl_step(corefl_ctx *c, double *dt_inout) {
  if (!c || !dt_inout) return static_cast<int>(Status::Arg);
  int st;
  if ((st = check_state_flag(c, false))) return st;
  if ((st = run_step(c, dt_inout))) return st;
  return post_state_flag(c);
}

int corefl_sync(corefl_ctx *c) {
  if (!c) return static_cast<int>(Status::Arg);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipGetLastError());
  return check_state_flag(c, true);
}

}  // extern "C"

namespace corefl {
namespace {

// One step of Wu's operator splitting (wu_splitting's loop body, src/StrangSplitting.cu:748-761): Cn
// (compute_cn, 15-41), the reaction sub-step with its boundary conditions, exchange and properties
// (wu_reaction_subStep, 479-506), then the RK3 sub-step over dt / 2 with dq - Cn and no property update
// (wu_RK_subStep, 568-638).
int wu_step(corefl_ctx *c, double *dt_inout) {
  int st;
  double dt = *dt_inout;
  if (!c->desc.fixed_time_step) {
    if ((st = global_dt(c, &dt))) return st;
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, Status::State, "non-positive or non-finite CFL time step");
  }
  if (c->run.hybrid && (st = op_all(c, COREFL_OP_SHOCK_SENSOR, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_ZERO_DQ, 0, 0))) return st;
  if ((st = op_all(c, COREFL_OP_VISCOUS, 0, 0))) return st;
  for (auto &b : c->blocks) {
    if ((st = op_block(c, COREFL_OP_INVISCID_X, b, 0, 0))) return st;
    if ((st = op_block(c, COREFL_OP_INVISCID_Y, b, 0, 0))) return st;
    if ((st = op_block(c, COREFL_OP_INVISCID_Z, b, 0, 0))) return st;
    if ((st = op_block(c, COREFL_OP_WU_CN, b, 0, 0))) return st;
  }
  for (auto &b : c->blocks) {
    if ((st = op_block(c, COREFL_OP_WU_REACTION, b, 0, dt))) return st;
    if ((st = op_block(c, COREFL_OP_BC, b, 0, dt))) return st;
  }
  if ((st = op_all(c, COREFL_OP_HALO, 0, dt))) return st;
  if ((st = op_all(c, COREFL_OP_K_GRADIENT, 0, dt))) return st;
  if ((st = op_all(c, COREFL_OP_PROPERTIES, 0, dt))) return st;
  const double dh = 0.5 * dt;
  if ((st = op_all(c, COREFL_OP_COPY_QN, 0, 0))) return st;
  for (int rk = 0; rk < 3; ++rk) {
    if ((st = op_all(c, COREFL_OP_ZERO_DQ, rk, dh))) return st;
    if ((st = op_all(c, COREFL_OP_VISCOUS, rk, dh))) return st;
    for (auto &b : c->blocks) {
      if ((st = op_block(c, COREFL_OP_INVISCID_X, b, rk, dh))) return st;
      if ((st = op_block(c, COREFL_OP_INVISCID_Y, b, rk, dh))) return st;
      if ((st = op_block(c, COREFL_OP_INVISCID_Z, b, rk, dh))) return st;
      if ((st = op_block(c, COREFL_OP_UPDATE_RK_WU, b, rk, dh))) return st;
      if (c->run.limit_flow && (st = op_block(c, COREFL_OP_LIMIT_FLOW, b, rk, dh))) return st;
      if ((st = op_block(c, COREFL_OP_BC, b, rk, dh))) return st;
    }
    if ((st = op_all(c, COREFL_OP_HALO, rk, dh))) return st;
    if ((st = op_all(c, COREFL_OP_K_GRADIENT, rk, dh))) return st;
  }
  *dt_inout = dt;
  return 0;
}

int run_stage(corefl_ctx *c, int rk, double dt) {
  int st;
  if ((st = op_all(c, COREFL_OP_ZERO_DQ, rk, dt))) return st;
  if ((st = op_all(c, COREFL_OP_VISCOUS, rk, dt))) return st;
  for (auto &b : c->blocks) {
    if ((st = op_block(c, COREFL_OP_CHEMISTRY, b, rk, dt))) return st;
    if ((st = op_block(c, COREFL_OP_INVISCID_X, b, rk, dt))) return st;
    if ((st = op_block(c, COREFL_OP_INVISCID_Y, b, rk, dt))) return st;
    if ((st = op_block(c, COREFL_OP_INVISCID_Z, b, rk, dt))) return st;
    if ((st = op_block(c, COREFL_OP_UPDATE_RK, b, rk, dt))) return st;
    if (c->run.limit_flow && (st = op_block(c, COREFL_OP_LIMIT_FLOW, b, rk, dt))) return st;
    if ((st = op_block(c, COREFL_OP_BC, b, rk, dt))) return st;
  }
  if ((st = op_all(c, COREFL_OP_HALO, rk, dt))) return st;
  if ((st = op_all(c, COREFL_OP_K_GRADIENT, rk, dt))) return st;
  if ((st = op_all(c, COREFL_OP_PROPERTIES, rk, dt))) return st;
  return 0;
}

int run_step(corefl_ctx *c, double *dt_inout) {
  int st;
  if (c->run.wu) return wu_step(c, dt_inout);
  if (c->desc.step_graph) {
    // one replay of the recorded step; the CFL dt comes back from device memory
    const bool fixed = c->desc.fixed_time_step != 0;
    if (c->blocks.empty()) return fail(c, Status::Arg, "no blocks");
    if (!c->graph_exec || (fixed && *dt_inout != c->graph_dt))
      if ((st = build_graph(c, *dt_inout, fixed))) return st;
    HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream));
    if (!fixed) {
      double dt = 0;
      HIP_TRY(c, hipMemcpyAsync(&dt, c->d_scalar + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(c, hipStreamSynchronize(c->stream));
      if (!(dt > 0) || !std::isfinite(dt)) return fail(c, Status::State, "non-positive or non-finite CFL time step");
      *dt_inout = dt;
    }
    return 0;
  }
  if ((st = op_all(c, COREFL_OP_COPY_QN, 0, 0))) return st;
  // compute_shock_sensor once per step for the hybrid WENO, src/RK.cuh:138-143
  if (c->run.hybrid && (st = op_all(c, COREFL_OP_SHOCK_SENSOR, 0, 0))) return st;
  double dt = *dt_inout;
  if (!c->desc.fixed_time_step) {
    if ((st = global_dt(c, &dt))) return st;
    if (!(dt > 0) || !std::isfinite(dt)) return fail(c, Status::State, "non-positive or non-finite CFL time step");
  }
  for (int rk = 0; rk < 3; ++rk)
    if ((st = run_stage(c, rk, dt))) return st;
  *dt_inout = dt;
  return 0;
}

}  // namespace
}  // namespace corefl

extern "C" {

int corefl_profile_begin(corefl_ctx *c) {
  if (!c) return static_cast<int>(Status::Arg);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->events.clear();
  c->pool_used = 0;
  c->profiling = true;
  return 0;
}

int corefl_profile_end(corefl_ctx *c, double *ms, long *calls) {
  if (!c) return static_cast<int>(Status::Arg);
  c->profiling = false;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (ms)
    for (int i = 0; i < COREFL_OP_COUNT; ++i) ms[i] = 0;
  if (calls)
    for (int i = 0; i < COREFL_OP_COUNT; ++i) calls[i] = 0;
  for (const auto &ev : c->events) {
    float t = 0;
    HIP_TRY(c, hipEventElapsedTime(&t, ev.a, ev.b));
    if (ms) ms[ev.op] += t;
    if (calls) calls[ev.op] += 1;
  }
  c->events.clear();
  c->pool_used = 0;
  return 0;
}

}  // extern "C"

namespace corefl {

CtxView ctx_view(const corefl_ctx *c) {
  CtxView v{&c->run, &c->host_tables, {}};
  for (const auto &b : c->blocks) v.blocks.push_back({b.z.mx, b.z.my, b.z.mz, b.z.ngg});
  return v;
}

int ctx_fail(corefl_ctx *c, Status s, const std::string &msg) { return fail(c, s, msg); }

}  // namespace corefl
