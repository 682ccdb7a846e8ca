// Allocation-free batch of served 0-D reactors (corefl_serve_run_noalloc).
// No counterpart in the reference; for callers that cannot allocate device
// memory (HIP graph capture, tight serving loops, harnesses that forbid it).
// It runs the reactor server's kernels (admission, then k_reactors until
// every request has reached t_end or failed) on one batch, with the chemistry
// tables, the slots and the admission records in the file-scope device array
// `scratch` below instead of memory from corefl_serve_open.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "corefl_internal.h"

namespace {

using corefl::Status;

constexpr size_t kScratchBytes = size_t(3) << 20;  // 3 MiB per GPU, held until the process ends
constexpr int kMaxDevices = 64;

// Device-side working memory of corefl_serve_run_noalloc: [Tab][slot arrays][admission records].
// Host code never passes this name to a kernel or a copy: it uses the device address that
// hipGetSymbolAddress gives for the current device (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per device: the address of `scratch` on that device (null until looked up).
unsigned char *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_lookup;

// Host staging, static as well: the derived tables, the admission records and the read-back.
corefl::Tab g_tab;
double g_rec[COREFL_NOALLOC_MAX_REACTORS * corefl::kReactorRecord];
int g_status[COREFL_NOALLOC_MAX_REACTORS];
double g_t[COREFL_NOALLOC_MAX_REACTORS], g_time[COREFL_NOALLOC_MAX_REACTORS], g_rho[COREFL_NOALLOC_MAX_REACTORS];
double g_rhoY[COREFL_MAX_SPEC * COREFL_NOALLOC_MAX_REACTORS];
long g_steps[COREFL_NOALLOC_MAX_REACTORS], g_rejected[COREFL_NOALLOC_MAX_REACTORS];

// Status codes only: a message would need storage this entry point does not take.
int nfail(Status st, const char *) { return static_cast<int>(st); }

#define NA_HIP(expr)                                      \
  do {                                                    \
    if ((expr) != hipSuccess) return nfail(Status::Hip, #expr); \
  } while (0)

// The device address of `scratch` on the current device, looked up once per device.
int scratch_ptr(unsigned char **out) {
  int dev = 0;
  NA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return nfail(Status::Arg, "device index beyond the supported range");
  std::lock_guard<std::mutex> lock(g_lookup);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    NA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Carves the slot arrays for n slots out of the scratch after the tables; returns the bytes used.
size_t carve(unsigned char *base, int n, int ns, corefl::ReactorPool &P, double **records) {
  size_t off = align256(sizeof(corefl::Tab));
  auto take = [&](auto *&p, size_t count) {
    p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + off);
    off = align256(off + count * sizeof(*p));
  };
  P.cap = n;
  take(P.rhoY, static_cast<size_t>(n) * ns);
  for (double **a : {&P.rho, &P.e, &P.t, &P.time, &P.t_end, &P.h, &P.err_last, &P.rtol, &P.atol}) take(*a, n);
  take(P.steps, n);
  take(P.rejected, n);
  take(P.max_steps, n);
  for (int **a : {&P.scheme, &P.adaptive, &P.conserve, &P.status}) take(*a, n);
  take(*records, static_cast<size_t>(n) * corefl::kReactorRecord);
  return off;
}

}  // namespace

extern "C" {

int corefl_serve_run_noalloc(const corefl_chem_tables *tables, const corefl_reactor_request *req, int n,
                             int steps_per_launch, corefl_reactor_result *res) {
  if (!tables || (n > 0 && (!req || !res))) return nfail(Status::Arg, "null argument");
  if (n < 0 || n > COREFL_NOALLOC_MAX_REACTORS)
    return nfail(Status::Arg, "the allocation-free reactor batch holds 0 to COREFL_NOALLOC_MAX_REACTORS requests");
  if (steps_per_launch < 1) return nfail(Status::Arg, "steps_per_launch must be positive");
  if (tables->n_spec < 1 || tables->n_spec > COREFL_MAX_SPEC || tables->n_reac < 0 ||
      tables->n_reac > COREFL_MAX_REAC)
    return nfail(Status::Arg, "table sizes exceed COREFL_MAX_SPEC / COREFL_MAX_REAC");
  if (tables->two_temperature) return nfail(Status::Arg, "the reactors run one-temperature mixtures");
  const int ns = tables->n_spec;
  const int model = ns | (tables->nasa7or9 == 9 ? 32 : 0);
  if (!corefl::model_supported(model) || model == 101)
    return nfail(Status::Arg, "no reactor kernel compiled for this mixture (5 or 9 species NASA-7, 5 NASA-9)");
  // the requests, checked as corefl_serve_submit checks them
  for (int i = 0; i < n; ++i) {
    const corefl_reactor_request &q = req[i];
    if (!(q.temperature > 0) || !(q.pressure > 0) || !(q.t_end > 0) || !(q.dt > 0))
      return nfail(Status::Arg, "temperature, pressure, t_end and dt must be positive");
    if (q.scheme != 0 && q.scheme != 1 && q.scheme != 3)
      return nfail(Status::Arg, "scheme must be 0 (point-implicit), 1 (ROK4E) or 3 (SSP-RK3)");
    if (q.scheme == 1 && q.adaptive && !(q.rtol > 0 && q.atol > 0))
      return nfail(Status::Arg, "adaptive ROK4E needs positive rtol and atol");
    double sum = 0;
    for (int l = 0; l < ns; ++l) {
      if (!(q.y[l] >= 0)) return nfail(Status::Arg, "mass fractions must be non-negative");
      sum += q.y[l];
    }
    if (std::fabs(sum - 1.0) > 1e-8) return nfail(Status::Arg, "mass fractions must sum to 1");
  }
  if (n == 0) return 0;
  unsigned char *base = nullptr;
  if (int st = scratch_ptr(&base)) return st;
  corefl::ReactorPool P{};
  double *d_rec = nullptr;
  if (carve(base, n, ns, P, &d_rec) > kScratchBytes) return nfail(Status::Arg, "the batch does not fit the scratch array");
  auto *d_tab = reinterpret_cast<corefl::Tab *>(base);
  // tables and admission records (host statics -> the scratch, on the null stream)
  corefl::derive_tables(*tables, g_tab);
  NA_HIP(hipMemcpy(d_tab, &g_tab, sizeof(corefl::Tab), hipMemcpyHostToDevice));
  const int R = corefl::kReactorRecord;
  for (int i = 0; i < n; ++i) {
    const corefl_reactor_request &q = req[i];
    double *rec = g_rec + static_cast<size_t>(i) * R;
    std::memset(rec, 0, R * sizeof(double));
    rec[0] = i;
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
    for (int l = 0; l < ns; ++l) rec[11 + l] = q.y[l];
  }
  NA_HIP(hipMemcpy(d_rec, g_rec, static_cast<size_t>(n) * R * sizeof(double), hipMemcpyHostToDevice));
  // every slot's fields, its status included, are written by the admission: nothing stale is read
  NA_HIP(corefl::launch_reactor_admit(P, model, d_tab, d_rec, n, nullptr));
  // the server's kernel choice: one wave per slot while few slots are in flight
  for (int active = n; active > 0;) {
    if (active <= corefl::kReactorWaveSlots)
      NA_HIP(corefl::launch_reactors_wave(P, model, d_tab, steps_per_launch, n, nullptr));
    else
      NA_HIP(corefl::launch_reactors(P, model, d_tab, steps_per_launch, nullptr));
    NA_HIP(hipMemcpy(g_status, P.status, n * sizeof(int), hipMemcpyDeviceToHost));
    active = 0;
    for (int i = 0; i < n; ++i) active += g_status[i] == 1;
  }
  NA_HIP(hipMemcpy(g_t, P.t, n * sizeof(double), hipMemcpyDeviceToHost));
  NA_HIP(hipMemcpy(g_time, P.time, n * sizeof(double), hipMemcpyDeviceToHost));
  NA_HIP(hipMemcpy(g_rho, P.rho, n * sizeof(double), hipMemcpyDeviceToHost));
  NA_HIP(hipMemcpy(g_steps, P.steps, n * sizeof(long), hipMemcpyDeviceToHost));
  NA_HIP(hipMemcpy(g_rejected, P.rejected, n * sizeof(long), hipMemcpyDeviceToHost));
  NA_HIP(hipMemcpy(g_rhoY, P.rhoY, static_cast<size_t>(n) * ns * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    corefl_reactor_result r{};
    const double rho = g_rho[i];
    double imw = 0;
    for (int l = 0; l < ns; ++l) {
      r.y[l] = g_rhoY[static_cast<size_t>(l) * n + i] / rho;
      imw += r.y[l] * tables->imw[l];
    }
    r.temperature = g_t[i];
    r.pressure = rho * corefl::kRu * imw * r.temperature;
    r.time = g_time[i];
    r.steps = g_steps[i];
    r.rejected = g_rejected[i];
    if (g_status[i] > corefl::kSlotFailed) {
      r.status = static_cast<int>(Status::State);
      r.reason = g_status[i] - corefl::kSlotFailed;
    }
    res[i] = r;
  }
  return 0;
}

}  // extern "C"
