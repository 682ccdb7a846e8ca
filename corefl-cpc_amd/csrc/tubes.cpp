// Request server for 1-D reactive shock tubes (no counterpart in the
// reference, which advances one grid per process).  Every tube in flight
// lives in one device-resident padded row: a tube of n nodes takes a stretch
// of n + 2 ngg row points (its nodes and its own ghost layers), found first-
// fit among the free stretches.  One RK3 step of all tubes is one launch per
// operator over the row (the order of corefl_stage: chemistry, x sweep,
// update, boundary conditions, properties), with each tube's CFL step and
// clock kept in its slot; a progress call replays its steps as one graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "corefl_internal.h"

using corefl::Status;

struct corefl_tube_server {
  int device = 0;
  hipStream_t stream = nullptr;
  corefl::Tab *d_tables = nullptr;
  corefl_chem_tables host{};
  corefl::Run run{};
  corefl::ChemJit jit;
  int g = 3, steps_per_progress = 1;
  long cap = 0;  // interior points of the row; the row holds cap + 2 g
  corefl::Zone z{};
  corefl::TubePool pool{};
  int *d_map = nullptr, *d_ranges = nullptr;
  double *d_dtc = nullptr, *d_adm = nullptr, *h_adm = nullptr;
  double *d_dth = nullptr;  // Wu's splitting: half of each point's step (its RK sub-step)
  double fixed_dt = 0.0;    // > 0: every tube steps by it (fixed_time_step), else its CFL step
  int *h_ranges = nullptr;
  std::vector<void *> allocs;
  // host view of the slots
  int *h_status = nullptr;
  double *h_time = nullptr, *h_profile = nullptr;
  long *h_steps = nullptr;
  std::vector<long> slot_ticket;  // -1: free
  std::vector<int> slot_seg, slot_n;  // row stretch (first row point) and node count
  std::vector<int> free_slots;
  int slot_high = 0;
  std::map<long, long> free_seg;     // row point -> length of a free stretch
  std::vector<std::pair<int, int>> to_clear;
  std::deque<std::pair<long, corefl_tube_request>> queue;
  struct Done {
    corefl_tube_result res{};
    std::vector<double> profile;
  };
  std::unordered_map<long, Done> done;
  long next_ticket = 0, in_flight = 0;
  // a progress call's steps recorded once as a graph, for the row extent and slot count it was built for
  hipGraphExec_t graph = nullptr;
  int graph_mx = -1, graph_slots = -1;
  std::string err;
};

namespace {

int tfail(corefl_tube_server *s, Status st, const std::string &m) {
  s->err = m;
  return static_cast<int>(st);
}

#define THIP(srv, expr)                                                                               \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return tfail(srv, Status::Hip, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(corefl_tube_server *s, T *&p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return e;
  s->allocs.push_back(q);
  p = static_cast<T *>(q);
  return hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T));
}

template <class T>
hipError_t halloc(T *&p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipHostMalloc(&q, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault);
  p = static_cast<T *>(q);
  return e;
}

constexpr int kMinTube = 16;

// first-fit stretch of len row points; -1 when none is free
long take_stretch(corefl_tube_server *s, long len) {
  for (auto it = s->free_seg.begin(); it != s->free_seg.end(); ++it) {
    if (it->second < len) continue;
    const long at = it->first, rest = it->second - len;
    s->free_seg.erase(it);
    if (rest > 0) s->free_seg[at + len] = rest;
    return at;
  }
  return -1;
}

void give_stretch(corefl_tube_server *s, long at, long len) {
  auto next = s->free_seg.lower_bound(at);
  if (next != s->free_seg.end() && next->first == at + len) {
    len += next->second;
    next = s->free_seg.erase(next);
  }
  if (next != s->free_seg.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == at) {
      prev->second += len;
      return;
    }
  }
  s->free_seg[at] = len;
}

// interior points the launches cover: up to the end of the last stretch in use
long row_extent(const corefl_tube_server *s) {
  long end = 0;
  for (int i = 0; i < s->slot_high; ++i)
    if (s->slot_ticket[i] >= 0) end = std::max<long>(end, s->slot_seg[i] + s->slot_n[i] + 2L * s->g);
  return std::max<long>(end - 2L * s->g, 0);
}

// One step of Wu's operator splitting for every tube in flight (wu_splitting's step,
// src/StrangSplitting.cu:748-761, as a one-row block runs it): each tube's step (fixed, or the CFL step
// without the chemical bound, local_time_step_without_reaction, src/TimeAdvanceFunc.cu:21-112); Cn from
// the x sweep; the reaction sub-step with each point's tube step, the tube ends and the properties; then
// the RK3 sub-step over half the step with dq - Cn and the tube ends after every stage.
int step_all_wu(corefl_tube_server *s, const corefl::Zone &z) {
  using namespace corefl;
  const Run &r = s->run;
  hipStream_t st = s->stream;
  const long len = z.mx + 2L * z.ngg;
  Run rr = r;
  rr.reaction = 0;
  THIP(s, launch_local_dt(z, rr, s->d_tables, s->d_dtc + s->cap, st));
  THIP(s, launch_tube_dt(z, s->pool, s->slot_high, st, s->fixed_dt));
  THIP(s, launch_tube_dtc(z, s->pool, s->d_map, s->d_dtc, st));
  THIP(s, launch_scaled_copy(s->d_dth, s->d_dtc, 0.5, z.mx, st));
  THIP(s, launch_fill(z.dq, z.dsz * r.nv, 0.0, st));
  THIP(s, launch_inviscid(z, r, s->d_tables, 0, st));
  THIP(s, launch_copy(z.cn, z.dq, z.dsz * r.nv, st));
  if (r.reaction) {
    if (s->jit.wu && r.chem_advance == 0)
      THIP(s, chem_jit_launch_wu(s->jit, z, 0.0, s->d_dtc, st));
    else
      THIP(s, launch_wu_reaction(z, r, s->d_tables, 0.0, r.chem_advance, st, s->d_dtc));
  }
  THIP(s, launch_tube_bc(z, s->pool, r.model, s->d_tables, s->slot_high, st));
  THIP(s, launch_properties_row(z, r, s->d_tables, st));
  THIP(s, launch_copy_planes(z.qn, z.cv, r.nv, z.sz, len, st));
  for (int rk = 0; rk < 3; ++rk) {
    THIP(s, launch_fill(z.dq, z.dsz * r.nv, 0.0, st));
    THIP(s, launch_inviscid(z, r, s->d_tables, 0, st));
    THIP(s, launch_update_rk_wu(z, r, s->d_tables, 0.0, rk, st, s->d_dth));
    THIP(s, launch_tube_bc(z, s->pool, r.model, s->d_tables, s->slot_high, st));
  }
  THIP(s, launch_tube_finish(s->pool, s->slot_high, st));
  return 0;
}

int step_all(corefl_tube_server *s, const corefl::Zone &z) {
  using namespace corefl;
  const Run &r = s->run;
  if (r.wu) return step_all_wu(s, z);
  hipStream_t st = s->stream;
  const long len = z.mx + 2L * z.ngg;
  THIP(s, launch_copy_planes(z.qn, z.cv, r.nv, z.sz, len, st));
  // dt_local of every point; the row's minimum goes to a spare word (each tube takes its own)
  THIP(s, launch_local_dt(z, r, s->d_tables, s->d_dtc + s->cap, st));
  THIP(s, launch_tube_dt(z, s->pool, s->slot_high, st, s->fixed_dt));
  THIP(s, launch_tube_dtc(z, s->pool, s->d_map, s->d_dtc, st));
  for (int rk = 0; rk < 3; ++rk) {
    THIP(s, launch_fill(z.dq, z.dsz * r.nv, 0.0, st));
    if (r.reaction) {
      if (s->jit.fn)
        THIP(s, chem_jit_launch(s->jit, z, st));
      else
        THIP(s, launch_chemistry(z, r, s->d_tables, st));
    }
    THIP(s, launch_inviscid(z, r, s->d_tables, 0, st));
    THIP(s, launch_update_rk(z, r, s->d_tables, 0.0, rk, st, nullptr, s->d_dtc));
    THIP(s, launch_tube_bc(z, s->pool, r.model, s->d_tables, s->slot_high, st));
    THIP(s, launch_properties_row(z, r, s->d_tables, st));
  }
  THIP(s, launch_tube_finish(s->pool, s->slot_high, st));
  return 0;
}

// The steps of one progress call as a graph: every kernel reads the tubes' steps and clocks from device
// memory, so the recording holds while the row extent and the slot count stay the same.
int record_steps(corefl_tube_server *s, const corefl::Zone &z) {
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  s->graph = nullptr;
  THIP(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int st = 0;
  for (int k = 0; k < s->steps_per_progress && !st; ++k) st = step_all(s, z);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s->stream, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  THIP(s, e);
  const hipError_t ei = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  THIP(s, ei);
  s->graph_mx = z.mx;
  s->graph_slots = s->slot_high;
  return 0;
}

}  // namespace

extern "C" {

int corefl_tubes_open(const corefl_run_desc *run, const corefl_chem_tables *tables, size_t budget_bytes,
                      int steps_per_progress, corefl_tube_server **out) {
  if (!out) return static_cast<int>(Status::Arg);
  auto *s = new corefl_tube_server();
  *out = s;
  if (!run || !tables) return tfail(s, Status::Arg, "the tube server needs a run description and chemistry tables");
  if (!run->mixture || tables->n_spec != run->n_spec || tables->n_spec < 1 || tables->n_spec > COREFL_MAX_SPEC ||
      tables->n_reac < 0 || tables->n_reac > COREFL_MAX_REAC)
    return tfail(s, Status::Arg, "the tube server runs a species mixture matching the tables");
  if (run->two_temperature || tables->two_temperature)
    return tfail(s, Status::Arg, "the tube server runs one-temperature mixtures");
  const int sch = run->inviscid_scheme;
  if (sch != 51 && sch != 52 && sch != 71 && sch != 72)
    return tfail(s, Status::Arg, "tubes use inviscid_scheme 51, 52, 71 or 72 (WENO)");
  if (run->viscous_order != 0 || run->positive_preserving || run->sensor_threshold > 1e-10 || run->limit_flow)
    return tfail(s, Status::Arg, "tubes are inviscid, without shock sensor, positivity limiter or limit_flow");
  if (!run->fixed_time_step && !(run->cfl > 0)) return tfail(s, Status::Arg, "tubes take the CFL step: cfl must be positive");
  if (run->fixed_time_step && !(run->dt > 0)) return tfail(s, Status::Arg, "a fixed time step must be positive");
  const int tsch = run->temporal_scheme == 0 ? 3 : run->temporal_scheme;
  if (tsch != 3 && tsch != 4) return tfail(s, Status::Arg, "temporal_scheme must be 3 (SSP-RK3) or 4 (Wu's splitting)");
  if (tsch == 4 && !run->reaction) return tfail(s, Status::Arg, "Wu's operator splitting needs reaction");
  if (run->chem_advance_method != 0 && run->chem_advance_method != 1 && run->chem_advance_method != 3)
    return tfail(s, Status::Arg, "chem_advance_method must be 0, 1 or 3");
  s->fixed_dt = run->fixed_time_step ? run->dt : 0.0;
  if (run->n_var != 5 + run->n_spec || run->n_scalar != run->n_spec)
    return tfail(s, Status::Arg, "n_var/n_scalar must be 5+n_spec/n_spec");
  if (steps_per_progress < 1) return tfail(s, Status::Arg, "steps_per_progress must be positive");
  s->steps_per_progress = steps_per_progress;
  s->host = *tables;
  corefl::Run &r = s->run;
  r.mixture = 1;
  r.ns = run->n_spec;
  r.model = r.ns | (tables->nasa7or9 == 9 ? 32 : 0);
  if (r.model != 5 && r.model != 9 && r.model != 37)
    return tfail(s, Status::Arg, "no tube kernels for this mixture (5 or 9 species NASA-7, 5 NASA-9)");
  r.nv = run->n_var;
  r.n_scalar = r.nst = run->n_spec;
  r.scheme = sch;
  r.reaction = run->reaction;
  r.cfl = run->cfl;
  r.prandtl = run->prandtl;
  r.eps_ref = 1e-40 * (run->rho_ref * run->v_ref * run->rho_ref * run->v_ref);
  r.rho_ref = run->rho_ref;
  r.v_ref = run->v_ref;
  r.wu = tsch == 4 ? 1 : 0;
  r.chem_advance = run->chem_advance_method;
  s->g = (sch == 71 || sch == 72) ? 4 : 3;
  const int nv = r.nv, ns = r.ns, g = s->g;
  // bytes per row point: cv, qn, bv, sv, metric, jac, gamma, c, Mach, reaction time scale (row);
  // dq, dt_local, the point's step and its slot (interior); admission and slot shares
  const size_t per_point = (2 * nv + 6 + ns + 9 + 4 + nv + 2 + (tsch == 4 ? nv + 1 : 0)) * sizeof(double) + sizeof(int) +
                           (corefl::kTubeRecord * sizeof(double) + 64) / kMinTube;
  const long cap = static_cast<long>(std::min<size_t>(budget_bytes / per_point, size_t(1) << 31)) - 2L * g;
  if (cap < kMinTube) return tfail(s, Status::Arg, "the budget holds no tube");
  s->cap = cap;
  const long row = cap + 2L * g;
  const int slots = static_cast<int>(std::max<long>(1, row / (kMinTube + 2L * g)));
  THIP(s, hipGetDevice(&s->device));
  THIP(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  THIP(s, hipMalloc(&s->d_tables, sizeof(corefl::Tab)));
  {
    auto t = std::make_unique<corefl::Tab>();
    corefl::derive_tables(s->host, *t);
    THIP(s, hipMemcpy(s->d_tables, t.get(), sizeof(corefl::Tab), hipMemcpyHostToDevice));
    if (r.reaction && run->chem_kernel == 0 && r.model == r.ns) {
      std::string e;
      const Status st = corefl::chem_jit_build(*t, r.ns, s->jit, e);
      if (st != Status::Ok) return tfail(s, st, "compiling the chemistry kernel for the mechanism: " + e);
    }
  }
  corefl::Zone &z = s->z;
  z.mx = static_cast<int>(cap);
  z.my = z.mz = 1;
  z.ngg = g;
  z.sx = z.sxy = 0;  // every (j, k) is the one row
  z.sz = row;
  z.off = g;
  z.dsz = cap;
  THIP(s, dalloc(s, z.cv, row * nv));
  THIP(s, dalloc(s, z.qn, row * nv));
  THIP(s, dalloc(s, z.bv, row * 6));
  THIP(s, dalloc(s, z.sv, row * ns));
  THIP(s, dalloc(s, z.metric, row * 9));
  THIP(s, dalloc(s, z.jac, row));
  THIP(s, dalloc(s, z.gam, row));
  THIP(s, dalloc(s, z.acoustic, row));
  THIP(s, dalloc(s, z.mach, row));
  THIP(s, dalloc(s, z.rts, row));
  THIP(s, dalloc(s, z.dq, cap * nv));
  // dt_local holds the per-block minima of launch_local_dt after its cap points
  THIP(s, dalloc(s, z.dt_local, cap + cap / 64 + 64));
  THIP(s, dalloc(s, s->d_dtc, cap + 1));
  if (r.wu) {
    THIP(s, dalloc(s, z.cn, cap * nv));
    THIP(s, dalloc(s, s->d_dth, cap));
  }
  THIP(s, dalloc(s, s->d_map, cap));
  THIP(s, hipMemset(s->d_map, 0xff, cap * sizeof(int)));  // -1: no tube
  // quiet filler state in the whole row (pure species 0 at rest, rho = 1, T = 300 K): points of no
  // tube stay finite under the row-wide launches
  {
    std::vector<double> h(row);
    auto put = [&](double *dst, double v) {
      std::fill(h.begin(), h.end(), v);
      return hipMemcpy(dst, h.data(), row * sizeof(double), hipMemcpyHostToDevice);
    };
    const double vals[6] = {1.0, 0.0, 0.0, 0.0, corefl::kRu * tables->imw[0] * 300.0, 300.0};
    for (int l = 0; l < 6; ++l) THIP(s, put(z.bv + l * row, vals[l]));
    THIP(s, put(z.sv, 1.0));
    for (int m = 0; m < 9; ++m) THIP(s, put(z.metric + m * row, (m == 0 || m == 4 || m == 8) ? 1.0 : 0.0));
    THIP(s, put(z.jac, 1.0));
    THIP(s, corefl::launch_cv_from_bv_row(z, r, s->d_tables, s->stream));
  }
  corefl::TubePool &P = s->pool;
  P.cap = slots;
  for (int **a : {&P.start, &P.n, &P.bc_l, &P.bc_r, &P.status, &P.last}) THIP(s, dalloc(s, *a, slots));
  for (double **a : {&P.time, &P.t_end, &P.dt}) THIP(s, dalloc(s, *a, slots));
  THIP(s, dalloc(s, P.steps, slots));
  THIP(s, dalloc(s, P.max_steps, slots));
  THIP(s, dalloc(s, s->d_adm, static_cast<size_t>(slots) * corefl::kTubeRecord));
  THIP(s, dalloc(s, s->d_ranges, 2 * static_cast<size_t>(slots)));
  THIP(s, halloc(s->h_adm, static_cast<size_t>(slots) * corefl::kTubeRecord));
  THIP(s, halloc(s->h_ranges, 2 * static_cast<size_t>(slots)));
  THIP(s, halloc(s->h_status, slots));
  THIP(s, halloc(s->h_time, slots));
  THIP(s, halloc(s->h_steps, slots));
  THIP(s, halloc(s->h_profile, static_cast<size_t>(6 + ns) * (cap + 2L * g)));
  // the zero fills of the allocations run on the null stream, which the server's non-blocking stream does
  // not wait for: finish them before any launch
  THIP(s, hipDeviceSynchronize());
  s->slot_ticket.assign(slots, -1);
  s->slot_seg.assign(slots, 0);
  s->slot_n.assign(slots, 0);
  for (int i = slots - 1; i >= 0; --i) s->free_slots.push_back(i);
  s->free_seg[0] = row;
  return 0;
}

long corefl_tubes_submit(corefl_tube_server *s, const corefl_tube_request *q) {
  if (!s) return static_cast<long>(Status::Arg);
  if (!q) return tfail(s, Status::Arg, "null request");
  if (q->n_cells < kMinTube) return tfail(s, Status::Arg, "a tube has at least 16 nodes");
  if (q->n_cells + 2L * s->g > s->cap + 2L * s->g) return tfail(s, Status::Arg, "the tube is longer than the row");
  if (q->n_left < 0 || q->n_left > q->n_cells) return tfail(s, Status::Arg, "n_left must lie in [0, n_cells]");
  if (!(q->length > 0) || !(q->t_end > 0)) return tfail(s, Status::Arg, "length and t_end must be positive");
  for (int b : {q->bc_left, q->bc_right})
    if (b != COREFL_BC_OUTFLOW && b != COREFL_BC_SYMMETRY)
      return tfail(s, Status::Arg, "tube ends are COREFL_BC_OUTFLOW or COREFL_BC_SYMMETRY");
  for (int side = 0; side < 2; ++side) {
    if (!(q->temperature[side] > 0) || !(q->pressure[side] > 0) || !std::isfinite(q->velocity[side]))
      return tfail(s, Status::Arg, "temperatures and pressures must be positive, velocities finite");
    double sum = 0;
    for (int l = 0; l < s->run.ns; ++l) {
      if (!(q->y[side][l] >= 0)) return tfail(s, Status::Arg, "mass fractions must be non-negative");
      sum += q->y[side][l];
    }
    if (std::fabs(sum - 1.0) > 1e-8) return tfail(s, Status::Arg, "mass fractions must sum to 1");
  }
  const long ticket = s->next_ticket++;
  s->queue.emplace_back(ticket, *q);
  return ticket;
}

int corefl_tubes_progress(corefl_tube_server *s, int *pending) {
  if (!s) return static_cast<int>(Status::Arg);
  using namespace corefl;
  const int g = s->g, R = kTubeRecord;
  // stretches of tubes retired at the last call leave the point -> slot map
  const int n_clear = static_cast<int>(s->to_clear.size());
  for (int i = 0; i < n_clear; ++i) {
    s->h_ranges[2 * i] = s->to_clear[i].first;
    s->h_ranges[2 * i + 1] = s->to_clear[i].second;
  }
  s->to_clear.clear();
  if (n_clear > 0) {
    THIP(s, hipMemcpyAsync(s->d_ranges, s->h_ranges, 2 * n_clear * sizeof(int), hipMemcpyHostToDevice, s->stream));
    THIP(s, launch_tube_clear(s->d_map, s->d_ranges, n_clear, s->cap, s->stream));
  }
  // admission in arrival order while the head of the queue finds a stretch and a slot
  int n_adm = 0;
  while (!s->queue.empty() && !s->free_slots.empty()) {
    const auto &[ticket, q] = s->queue.front();
    const long at = take_stretch(s, q.n_cells + 2L * g);
    if (at < 0) break;
    const int slot = s->free_slots.back();
    s->free_slots.pop_back();
    double *rec = s->h_adm + static_cast<size_t>(n_adm) * R;
    std::memset(rec, 0, R * sizeof(double));
    rec[0] = slot;
    rec[1] = static_cast<double>(at);  // interior index of the first node: row point at + g, minus g
    rec[2] = q.n_cells;
    rec[3] = q.n_left;
    // derived numbers in the order a one-block setup computes them: dx, 1 / dx, dx^3 and p / (R T)
    // with R summed in species order
    const double dx = q.length / (q.n_cells - 1);
    rec[4] = dx;
    rec[16] = 1.0 / dx;
    rec[17] = dx * dx * dx;
    rec[5] = q.t_end;
    rec[18] = static_cast<double>(q.max_steps > 0 ? q.max_steps : 0);
    rec[6] = q.bc_left;
    rec[7] = q.bc_right;
    for (int side = 0; side < 2; ++side) {
      rec[8 + side] = q.temperature[side];
      rec[10 + side] = q.pressure[side];
      rec[12 + side] = q.velocity[side];
      double R = 0.0;
      for (int l = 0; l < s->run.ns; ++l) {
        const double a = q.y[side][l] * s->host.gas_const[l];
        R = R + a;
      }
      rec[14 + side] = q.pressure[side] / (R * q.temperature[side]);
      for (int l = 0; l < s->run.ns; ++l) rec[19 + side * COREFL_MAX_SPEC + l] = q.y[side][l];
    }
    s->slot_ticket[slot] = ticket;
    s->slot_seg[slot] = static_cast<int>(at);
    s->slot_n[slot] = q.n_cells;
    s->slot_high = std::max(s->slot_high, slot + 1);
    ++s->in_flight;
    ++n_adm;
    s->queue.pop_front();
  }
  Zone z = s->z;
  z.mx = static_cast<int>(row_extent(s));
  z.dsz = z.mx;
  if (n_adm > 0) {
    THIP(s, hipMemcpyAsync(s->d_adm, s->h_adm, static_cast<size_t>(n_adm) * R * sizeof(double),
                           hipMemcpyHostToDevice, s->stream));
    THIP(s, launch_tube_admit(z, s->pool, s->run.model, s->d_tables, s->d_adm, n_adm, s->d_map, s->cap, s->stream));
    // the one-block initialisation: boundary conditions, then cv of the nodes, then properties
    // (the boundary kernels and the properties give the tubes in flight the values they hold)
    THIP(s, launch_tube_bc(z, s->pool, s->run.model, s->d_tables, s->slot_high, s->stream));
    THIP(s, launch_tube_cv(z, s->run.model, s->d_tables, s->d_adm, n_adm, s->stream));
    THIP(s, launch_properties_row(z, s->run, s->d_tables, s->stream));
  }
  if (s->in_flight > 0 && z.mx > 0) {
    if (!s->graph || s->graph_mx != z.mx || s->graph_slots != s->slot_high)
      if (int st = record_steps(s, z)) return st;
    THIP(s, hipGraphLaunch(s->graph, s->stream));
    const size_t n = s->slot_high;
    THIP(s, hipMemcpyAsync(s->h_status, s->pool.status, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    THIP(s, hipMemcpyAsync(s->h_time, s->pool.time, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    THIP(s, hipMemcpyAsync(s->h_steps, s->pool.steps, n * sizeof(long), hipMemcpyDeviceToHost, s->stream));
    THIP(s, hipStreamSynchronize(s->stream));
    const int ns = s->run.ns;
    for (int i = 0; i < s->slot_high; ++i) {
      const int st = s->h_status[i];
      if (s->slot_ticket[i] < 0 || (st != 2 && st <= corefl::kSlotFailed)) continue;
      const int nc = s->slot_n[i];
      const long first = s->slot_seg[i] + g;  // row point of node 0
      const size_t pitch = static_cast<size_t>(z.sz) * sizeof(double);
      THIP(s, hipMemcpy2DAsync(s->h_profile, nc * sizeof(double), z.bv + first, pitch, nc * sizeof(double), 6,
                               hipMemcpyDeviceToHost, s->stream));
      THIP(s, hipMemcpy2DAsync(s->h_profile + 6L * nc, nc * sizeof(double), z.sv + first, pitch,
                               nc * sizeof(double), ns, hipMemcpyDeviceToHost, s->stream));
      THIP(s, hipStreamSynchronize(s->stream));
      corefl_tube_server::Done d;
      d.res.time = s->h_time[i];
      d.res.steps = s->h_steps[i];
      d.res.n_cells = nc;
      if (st > corefl::kSlotFailed) {
        d.res.status = static_cast<int>(Status::State);
        d.res.reason = st - corefl::kSlotFailed;
      }
      d.profile.assign(s->h_profile, s->h_profile + static_cast<size_t>(6 + ns) * nc);
      s->done[s->slot_ticket[i]] = std::move(d);
      s->to_clear.emplace_back(s->slot_seg[i] - g, nc + 2 * g);
      give_stretch(s, s->slot_seg[i], nc + 2L * g);
      s->slot_ticket[i] = -1;
      s->free_slots.push_back(i);
      --s->in_flight;
    }
    while (s->slot_high > 0 && s->slot_ticket[s->slot_high - 1] < 0) --s->slot_high;
  }
  if (pending) *pending = static_cast<int>(s->in_flight + static_cast<long>(s->queue.size()));
  return 0;
}

int corefl_tubes_poll(corefl_tube_server *s, long ticket, corefl_tube_result *res, double *profile,
                      long profile_len) {
  if (!s) return static_cast<int>(Status::Arg);
  auto it = s->done.find(ticket);
  if (it != s->done.end()) {
    if (profile) {
      if (profile_len < static_cast<long>(it->second.profile.size()))
        return tfail(s, Status::Arg, "the profile buffer holds fewer than (6 + n_spec) n_cells values");
      std::memcpy(profile, it->second.profile.data(), it->second.profile.size() * sizeof(double));
    }
    if (res) *res = it->second.res;
    s->done.erase(it);
    return 1;
  }
  if (ticket >= 0 && ticket < s->next_ticket) return 0;
  return tfail(s, Status::Arg, "unknown ticket");
}

int corefl_tubes_ready(const corefl_tube_server *s, long *tickets, int max) {
  if (!s || max < 0 || (max > 0 && !tickets)) return static_cast<int>(Status::Arg);
  int n = 0;
  for (auto it = s->done.begin(); it != s->done.end() && n < max; ++it) tickets[n++] = it->first;
  return n;
}

long corefl_tubes_capacity(const corefl_tube_server *s) { return s ? s->cap + 2L * s->g : 0; }

const char *corefl_tubes_error(const corefl_tube_server *s) { return s ? s->err.c_str() : "null server"; }

void corefl_tubes_close(corefl_tube_server *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  for (void *p : s->allocs) (void)hipFree(p);
  for (void *p : {static_cast<void *>(s->h_adm), static_cast<void *>(s->h_ranges), static_cast<void *>(s->h_status),
                  static_cast<void *>(s->h_time), static_cast<void *>(s->h_steps),
                  static_cast<void *>(s->h_profile)})
    if (p) (void)hipHostFree(p);
  if (s->d_tables) (void)hipFree(s->d_tables);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // extern "C"
