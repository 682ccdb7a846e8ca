// Request server for 0-D constant-volume reactors: device-resident slots,
// continuous admission, one launch per progress call (no counterpart in the
// reference, whose const_volume_reactor integrates one case per run,
// src/ConstVolumeReactor.cpp:320-775).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "corefl_internal.h"

using corefl::Status;

struct corefl_server {
  int device = 0;
  hipStream_t stream = nullptr;
  corefl::Tab *d_tables = nullptr;
  corefl_chem_tables host{};
  int model = 0, ns = 0, steps_per_launch = 1;
  corefl::ReactorPool pool{};
  std::vector<void *> allocs;
  double *d_adm = nullptr, *h_adm = nullptr;  // admission records (device, pinned host)
  // pinned read-back of the slots in use
  int *h_status = nullptr;
  double *h_t = nullptr, *h_time = nullptr, *h_rho = nullptr, *h_rhoY = nullptr;
  long *h_steps = nullptr, *h_rejected = nullptr;
  std::vector<long> slot_ticket;  // -1: free
  std::vector<int> free_slots;
  int high_water = 0;             // slots [0, high_water) have been used
  std::deque<std::pair<long, corefl_reactor_request>> queue;
  std::unordered_map<long, corefl_reactor_result> done;
  long next_ticket = 0;
  long in_flight = 0;
  std::string err;
};

namespace {

int sfail(corefl_server *s, Status st, const std::string &m) {
  s->err = m;
  return static_cast<int>(st);
}

#define SHIP(srv, expr)                                                                               \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return sfail(srv, Status::Hip, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(corefl_server *s, T *&p, size_t n) {
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

}  // namespace

extern "C" {

int corefl_serve_open(const corefl_chem_tables *tables, size_t budget_bytes, int steps_per_launch,
                      corefl_server **out) {
  if (!out) return static_cast<int>(Status::Arg);
  auto *s = new corefl_server();
  *out = s;
  if (!tables) return sfail(s, Status::Arg, "the reactor server needs chemistry tables");
  if (tables->n_spec < 1 || tables->n_spec > COREFL_MAX_SPEC || tables->n_reac < 0 ||
      tables->n_reac > COREFL_MAX_REAC)
    return sfail(s, Status::Arg, "table sizes exceed COREFL_MAX_SPEC / COREFL_MAX_REAC");
  if (tables->two_temperature) return sfail(s, Status::Arg, "the reactor server runs one-temperature mixtures");
  s->host = *tables;
  s->ns = tables->n_spec;
  s->model = s->ns | (tables->nasa7or9 == 9 ? 32 : 0);
  if (!corefl::model_supported(s->model) || s->model == 101)
    return sfail(s, Status::Arg, "no reactor kernel compiled for this mixture (5 or 9 species NASA-7, 5 NASA-9)");
  if (steps_per_launch < 1) return sfail(s, Status::Arg, "steps_per_launch must be positive");
  s->steps_per_launch = steps_per_launch;
  const size_t per_slot = (static_cast<size_t>(s->ns) + 10) * sizeof(double) + 3 * sizeof(long) + 4 * sizeof(int) +
                          corefl::kReactorRecord * sizeof(double);
  const size_t cap = std::min<size_t>(budget_bytes / per_slot, 1u << 20);
  if (cap < 1) return sfail(s, Status::Arg, "the budget holds no reactor slot");
  corefl::ReactorPool &P = s->pool;
  P.cap = static_cast<int>(cap);
  SHIP(s, hipGetDevice(&s->device));
  SHIP(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  SHIP(s, hipMalloc(&s->d_tables, sizeof(corefl::Tab)));
  {
    auto t = std::make_unique<corefl::Tab>();
    corefl::derive_tables(s->host, *t);
    SHIP(s, hipMemcpy(s->d_tables, t.get(), sizeof(corefl::Tab), hipMemcpyHostToDevice));
  }
  SHIP(s, dalloc(s, P.rhoY, cap * s->ns));
  for (double **a : {&P.rho, &P.e, &P.t, &P.time, &P.t_end, &P.h, &P.err_last, &P.rtol, &P.atol})
    SHIP(s, dalloc(s, *a, cap));
  SHIP(s, dalloc(s, P.steps, cap));
  SHIP(s, dalloc(s, P.rejected, cap));
  SHIP(s, dalloc(s, P.max_steps, cap));
  for (int **a : {&P.scheme, &P.adaptive, &P.conserve, &P.status}) SHIP(s, dalloc(s, *a, cap));
  SHIP(s, dalloc(s, s->d_adm, cap * corefl::kReactorRecord));
  SHIP(s, halloc(s->h_adm, cap * corefl::kReactorRecord));
  SHIP(s, halloc(s->h_status, cap));
  SHIP(s, halloc(s->h_t, cap));
  SHIP(s, halloc(s->h_time, cap));
  SHIP(s, halloc(s->h_rho, cap));
  SHIP(s, halloc(s->h_rhoY, cap * s->ns));
  SHIP(s, halloc(s->h_steps, cap));
  SHIP(s, halloc(s->h_rejected, cap));
  // the zero fills of the allocations run on the null stream, which the server's non-blocking stream does
  // not wait for: finish them before any launch
  SHIP(s, hipDeviceSynchronize());
  s->slot_ticket.assign(cap, -1);
  s->free_slots.reserve(cap);
  for (int i = static_cast<int>(cap) - 1; i >= 0; --i) s->free_slots.push_back(i);
  return 0;
}

long corefl_serve_submit(corefl_server *s, const corefl_reactor_request *q) {
  if (!s) return static_cast<long>(Status::Arg);
  if (!q) return sfail(s, Status::Arg, "null request");
  if (!(q->temperature > 0) || !(q->pressure > 0) || !(q->t_end > 0) || !(q->dt > 0))
    return sfail(s, Status::Arg, "temperature, pressure, t_end and dt must be positive");
  if (q->scheme != 0 && q->scheme != 1 && q->scheme != 3)
    return sfail(s, Status::Arg, "scheme must be 0 (point-implicit), 1 (ROK4E) or 3 (SSP-RK3)");
  if (q->scheme == 1 && q->adaptive && !(q->rtol > 0 && q->atol > 0))
    return sfail(s, Status::Arg, "adaptive ROK4E needs positive rtol and atol");
  double sum = 0;
  for (int l = 0; l < s->ns; ++l) {
    if (!(q->y[l] >= 0)) return sfail(s, Status::Arg, "mass fractions must be non-negative");
    sum += q->y[l];
  }
  if (std::fabs(sum - 1.0) > 1e-8) return sfail(s, Status::Arg, "mass fractions must sum to 1");
  const long ticket = s->next_ticket++;
  s->queue.emplace_back(ticket, *q);
  return ticket;
}

int corefl_serve_progress(corefl_server *s, int *pending) {
  if (!s) return static_cast<int>(Status::Arg);
  corefl::ReactorPool &P = s->pool;
  const int R = corefl::kReactorRecord;
  // admission into free slots
  int n_adm = 0;
  while (!s->queue.empty() && !s->free_slots.empty()) {
    const int slot = s->free_slots.back();
    s->free_slots.pop_back();
    const auto &[ticket, q] = s->queue.front();
    double *rec = s->h_adm + static_cast<size_t>(n_adm) * R;
    std::memset(rec, 0, R * sizeof(double));
    rec[0] = slot;
    rec[1] = q.temperature;
    rec[2] = q.pressure;
    rec[3] = q.t_end;
    rec[4] = q.dt;
    rec[5] = q.rtol;
    rec[6] = q.atol;
    rec[7] = q.scheme;
    rec[8] = q.adaptive;
    rec[9] = q.conserve_mass;
    rec[10] = static_cast<double>(q.max_steps > 0 ? q.max_steps : 0);
    for (int l = 0; l < s->ns; ++l) rec[11 + l] = q.y[l];
    s->slot_ticket[slot] = ticket;
    s->high_water = std::max(s->high_water, slot + 1);
    ++s->in_flight;
    ++n_adm;
    s->queue.pop_front();
  }
  if (n_adm > 0) {
    SHIP(s, hipMemcpyAsync(s->d_adm, s->h_adm, static_cast<size_t>(n_adm) * R * sizeof(double),
                           hipMemcpyHostToDevice, s->stream));
    SHIP(s, corefl::launch_reactor_admit(P, s->model, s->d_tables, s->d_adm, n_adm, s->stream));
  }
  if (s->in_flight > 0) {
    // few slots in flight: one wave per slot (latency); many: one lane per slot (throughput).  Both give
    // the same results bit for bit, so a request's result does not depend on the load.
    if (s->in_flight <= corefl::kReactorWaveSlots)
      SHIP(s, corefl::launch_reactors_wave(P, s->model, s->d_tables, s->steps_per_launch, s->high_water, s->stream));
    else
      SHIP(s, corefl::launch_reactors(P, s->model, s->d_tables, s->steps_per_launch, s->stream));
    const size_t n = s->high_water;
    SHIP(s, hipMemcpyAsync(s->h_status, P.status, n * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(s->h_t, P.t, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(s->h_time, P.time, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(s->h_rho, P.rho, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(s->h_steps, P.steps, n * sizeof(long), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(s->h_rejected, P.rejected, n * sizeof(long), hipMemcpyDeviceToHost, s->stream));
    for (int l = 0; l < s->ns; ++l)
      SHIP(s, hipMemcpyAsync(s->h_rhoY + static_cast<size_t>(l) * P.cap, P.rhoY + static_cast<size_t>(l) * P.cap,
                             n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->high_water; ++i) {
      const int st = s->h_status[i];
      if (s->slot_ticket[i] < 0 || (st != 2 && st <= corefl::kSlotFailed)) continue;
      corefl_reactor_result r{};
      if (st > corefl::kSlotFailed) {
        r.status = static_cast<int>(Status::State);
        r.reason = st - corefl::kSlotFailed;
      }
      const double rho = s->h_rho[i];
      double imw = 0;
      for (int l = 0; l < s->ns; ++l) {
        r.y[l] = s->h_rhoY[static_cast<size_t>(l) * P.cap + i] / rho;
        imw += r.y[l] * s->host.imw[l];
      }
      r.temperature = s->h_t[i];
      r.pressure = rho * corefl::kRu * imw * r.temperature;
      r.time = s->h_time[i];
      r.steps = s->h_steps[i];
      r.rejected = s->h_rejected[i];
      s->done[s->slot_ticket[i]] = r;
      s->slot_ticket[i] = -1;
      s->free_slots.push_back(i);
      --s->in_flight;
    }
  }
  if (pending) *pending = static_cast<int>(s->in_flight + static_cast<long>(s->queue.size()));
  return 0;
}

int corefl_serve_poll(corefl_server *s, long ticket, corefl_reactor_result *res) {
  if (!s) return static_cast<int>(Status::Arg);
  auto it = s->done.find(ticket);
  if (it != s->done.end()) {
    if (res) *res = it->second;
    s->done.erase(it);
    return 1;
  }
  if (ticket >= 0 && ticket < s->next_ticket) return 0;
  return sfail(s, Status::Arg, "unknown ticket");
}

int corefl_serve_drain(corefl_server *s, long *tickets, corefl_reactor_result *res, int max) {
  if (!s || max < 0 || (max > 0 && (!tickets || !res))) return static_cast<int>(Status::Arg);
  int n = 0;
  for (auto it = s->done.begin(); it != s->done.end() && n < max;) {
    tickets[n] = it->first;
    res[n] = it->second;
    ++n;
    it = s->done.erase(it);
  }
  return n;
}

int corefl_serve_capacity(const corefl_server *s) { return s ? s->pool.cap : 0; }

const char *corefl_serve_error(const corefl_server *s) { return s ? s->err.c_str() : "null server"; }

void corefl_serve_close(corefl_server *s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (void *p : s->allocs) (void)hipFree(p);
  for (void *p : {static_cast<void *>(s->h_adm), static_cast<void *>(s->h_status), static_cast<void *>(s->h_t),
                  static_cast<void *>(s->h_time), static_cast<void *>(s->h_rho), static_cast<void *>(s->h_rhoY),
                  static_cast<void *>(s->h_steps), static_cast<void *>(s->h_rejected)})
    if (p) (void)hipHostFree(p);
  if (s->d_tables) (void)hipFree(s->d_tables);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // extern "C"
