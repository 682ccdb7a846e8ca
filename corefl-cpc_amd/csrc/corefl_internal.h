// Internal declarations shared by the host driver (driver.cpp, chemdata.cpp,
// setup.cpp) and the device code (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <variant>
#include <vector>

#include "corefl_amd.h"

namespace corefl {

enum class Status : int { Ok = 0, Arg = -1, Hip = -2, Io = -3, Parse = -4, State = -5 };

constexpr double kRu = 8314.462618;  // J/(kmol K), src/Constants.h:19

Status load_chemistry(const std::string &mech, const std::string &therm, const std::string &tran,
                      corefl_chem_tables &t, std::string &err);
Status load_chemistry_ex(const std::string &mech, const std::string &therm, const std::string &tran, int nasa,
                         const std::string &two_temperature, corefl_chem_tables &t, std::string &err);

// A setup.txt as the reference's Parameter reads it (src/Parameter.cpp:38-184).
using SetupStruct = std::map<std::string, std::variant<std::string, int, double>>;
struct Setup {
  std::map<std::string, int> ints;
  std::map<std::string, double> reals;
  std::map<std::string, bool> bools;
  std::map<std::string, std::string> strings;
  std::map<std::string, std::vector<std::string>> string_arrays;
  std::map<std::string, SetupStruct> structs;
};
bool read_setup(const std::string &path, Setup &s, std::string &err);
double struct_real(const SetupStruct &st, const std::string &k, double def);
// species l's name in the mechanism file the setup names (case_dir/input/ + mechanism_file)
std::string species_name(const corefl_chem_tables &t, const std::string &input_dir, const Setup &s, int l);
template <class M, class K, class V>
V get_or(const M &m, const K &k, V def) {
  auto it = m.find(k);
  return it == m.end() ? def : it->second;
}

// Device view of one block: padded SoA planes (i fastest) as in the
// reference's DZone (src/Field.h:12-112).  Passed to kernels by value.
struct Zone {
  int mx, my, mz, ngg;
  long sx, sxy, sz, off;  // padded strides, plane size, offset of (0,0,0)
  long dsz;               // interior plane size (dq, dt_local)
  double *cv, *qn, *bv, *sv, *dq, *ff, *gf, *hf, *metric, *jac;
  double *mul, *kappa, *gam, *acoustic, *mach, *rhoD, *rts, *dt_local;
  double *zg;  // zeta derivatives of u, v, w, T, Y_l (, p) for the 8th-order viscous fluxes (interior)
  double *tve, *kappa_ve;  // two-temperature: temperature_ve, thermal_conductivity_ve (padded)
  double *sensor;          // shock sensor (padded; hybrid WENO only, ghosts stay 0)
  int *bad;                // set to 1 by the update where a point's state is not finite and positive
  double *cn;              // Wu splitting: the transport part of dq frozen over a step (Cn, interior, nv planes)
  int *bt_il, *bt_ir, *bt_jl, *bt_jr, *bt_kl, *bt_kr;

  __host__ __device__ long ix(int i, int j, int k) const {
    return static_cast<long>(k) * sxy + static_cast<long>(j) * sx + i + off;
  }
  __host__ __device__ long dx(int i, int j, int k) const {
    return (static_cast<long>(k) * my + j) * static_cast<long>(mx) + i;
  }
};

// Request-server slots of 0-D constant-volume reactors, structure of arrays
// over the capacity (rhoY is [n_spec][cap]).  status: 0 free, 1 active, 2 done,
// kSlotFailed + COREFL_FAIL_* reason: failed.
struct ReactorPool {
  int cap;
  double *rhoY, *rho, *e, *t, *time, *t_end, *h, *err_last, *rtol, *atol;
  long *steps, *rejected, *max_steps;
  int *scheme, *adaptive, *conserve, *status;
};
constexpr int kSlotFailed = 8;  // status kSlotFailed + reason (1 non-finite, 2 step collapse, 3 step cap)
// one admission record: slot, T0, p0, t_end, h0, rtol, atol, scheme, adaptive, conserve, max_steps,
// Y[COREFL_MAX_SPEC]
constexpr int kReactorRecord = 11 + COREFL_MAX_SPEC;

// Request-server slots of 1-D shock tubes packed end to end in one padded
// row (each tube with its own ghost layers); start is the row's interior
// index of the tube's first node.  status: 0 free, 1 active, 2 done,
// kSlotFailed + reason: failed.
struct TubePool {
  int cap;
  int *start, *n, *bc_l, *bc_r, *status, *last;
  double *time, *t_end, *dt;
  long *steps, *max_steps;
};
// one admission record: slot, start, n, n_left, dx, t_end, bc_l, bc_r, T[2], p[2], u[2], rho[2], 1 / dx, jac,
// max_steps, Y_left[COREFL_MAX_SPEC], Y_right[COREFL_MAX_SPEC]
constexpr int kTubeRecord = 19 + 2 * COREFL_MAX_SPEC;

// Scalar switches the kernels read (subset of DParameter).
struct Run {
  int mixture, ns, nv, n_scalar, nst;  // nst = n_scalar_transported
  int model;  // kernel instantiation: species count | 32 (NASA-9) | 64 (two-temperature); 0 for Air
  int scheme, viscous_order, reaction, gradp;
  double prandtl, cfl, eps_ref;
  int recon, limiter;  // primitive-variable path: reconstruction 1/2/3, limiter 0/1
  double efix;         // Roe entropy-fix factor
  int limit_flow;
  double ll[5], ul[5];  // limit_flow bounds on rho, u, v, w, p
  int tt, vt;           // two-temperature model (E_ve is scalar ns), V-T relaxation after the update
  int pp;               // species positivity limiter of the WENO fluxes
  double pp_dt;         // DParameter::dt (fixed step) or 0
  bool hybrid;          // WENO scheme with shockSensor_threshold > 1e-10
  int sensor_type;      // 0 Ducros, 1 Jameson, 2 density / pressure jump
  double sensor_eps, sensor_threshold;
  double rho_ref, v_ref;
  int wu;               // temporal_scheme 4: Wu's operator splitting
  int chem_advance;     // its reaction sub-step: 0 explicit, 1 point-implicit, 3 SSP-RK3
};

// One reaction with its stoichiometry as sparse lists in ascending species
// order, so that sums over them visit the nonzero terms of the dense sums in
// the same order.
struct Reac {
  int nf, nb, nnu, npart;
  int f_idx[COREFL_MAX_SPEC], f_ord[COREFL_MAX_SPEC];  // reactants: species, exponent
  int b_idx[COREFL_MAX_SPEC], b_ord[COREFL_MAX_SPEC];  // products
  int nu_idx[COREFL_MAX_SPEC];                         // species with nu_b - nu_f != 0
  double nu[COREFL_MAX_SPEC];
  int part_idx[COREFL_MAX_SPEC];                       // species with nu_f != 0 or nu_b != 0
  double part_f[COREFL_MAX_SPEC], part_b[COREFL_MAX_SPEC];
};

// Device copy of the tables plus values derived once at context creation
// (constant factors of the transport fits, sparse reaction lists).
struct Tab : corefl_chem_tables {
  double pw_lj[COREFL_MAX_SPEC];                    // LJ_potent_inv^-0.145
  double pw_kbe[COREFL_MAX_SPEC][COREFL_MAX_SPEC];  // 1.0548 kb_over_eps_jk^-0.15504
  double inv_bdc[COREFL_MAX_SPEC][COREFL_MAX_SPEC]; // 1 / binary_diffusivity_coeff
  Reac reac[COREFL_MAX_REAC];
};
void derive_tables(const corefl_chem_tables &in, Tab &out);

// A context's run and blocks as host code beside the driver sees them (the
// solution-file reader and writer, plt.cpp), and its error slot.
struct BlockExtent {
  int mx, my, mz, ngg;
};
struct CtxView {
  const Run *run;
  const corefl_chem_tables *tables;
  std::vector<BlockExtent> blocks;
};
CtxView ctx_view(const corefl_ctx *c);
int ctx_fail(corefl_ctx *c, Status s, const std::string &msg);

// A chemistry kernel compiled for one mechanism (hipRTC), loaded on one device.
// Compiled modules are cached per (device, generated source) for the life of
// the process.
struct ChemJit {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;  // corefl_chem: source terms into dq (k_chemistry)
  hipFunction_t wu = nullptr;  // corefl_wu: Wu's explicit reaction step (k_wu_reaction, method 0)
};
Status chem_jit_build(const Tab &tb, int ns, ChemJit &out, std::string &err);
std::string chem_jit_source(const Tab &tb, int ns);
hipError_t chem_jit_launch(const ChemJit &j, const Zone &z, hipStream_t s);
// dtc: a step per interior point (batched tubes), else every point takes dt
hipError_t chem_jit_launch_wu(const ChemJit &j, const Zone &z, double dt, const double *dtc, hipStream_t s);

// A registered 1-to-1 interface face (InnerFace after register_boundary,
// src/Mesh.h:33-59, src/Mesh.cpp:39-172): the loop over the face node and its
// ghost layers and the index map into the target block (0-based).
struct InnerDev {
  int range_start[3], range_end[3], face, direction;
  int target_start[3], target_face, target_direction;
  int src_tar[3], loop_dir[3], target_loop_dir[3], n_point[3];
  int target_block;
};
void register_inner(const corefl_inner_face &in, int dim, int ngg, InnerDev &f);

// ---- kernel launchers (kernels.hip) ----
hipError_t launch_fill(double *p, long n, double v, hipStream_t s);
hipError_t launch_copy(double *dst, const double *src, long n, hipStream_t s);
hipError_t launch_scaled_copy(double *dst, const double *src, double a, long n, hipStream_t s);
hipError_t launch_cv_from_bv(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
hipError_t launch_properties(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
// save_qn: the first stage of a recorded step, whose cv is the step's start state, saves qn itself
hipError_t launch_update_rk(const Zone &z, const Run &r, const Tab *tb, double dt, int rk, hipStream_t s,
                            const double *dtp = nullptr, const double *dtc = nullptr, bool save_qn = false);
hipError_t launch_chemistry(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
// Wu's splitting: the reaction sub-step over dt (dtc: per point) and the RK sub-step's update with dq - Cn
hipError_t launch_wu_reaction(const Zone &z, const Run &r, const Tab *tb, double dt, int method, hipStream_t s,
                              const double *dtc = nullptr);
hipError_t launch_update_rk_wu(const Zone &z, const Run &r, const Tab *tb, double dt, int rk, hipStream_t s,
                               const double *dtc = nullptr);
hipError_t launch_vt_relaxation(const Zone &z, const Run &r, const Tab *tb, double dt_stage, hipStream_t s,
                                const double *dtp = nullptr, double scale = 1.0);
hipError_t launch_min_combine(const double *v, int n, double *out, hipStream_t s);
// shape: tile shape of the characteristic WENO sweep (0 or 1, kWenoShapes of them; same results)
constexpr int kWenoShapes = 2;
hipError_t launch_inviscid(const Zone &z, const Run &r, const Tab *tb, int dir, hipStream_t s, int shape = 0);
hipError_t launch_shock_sensor(const Zone &z, const Run &r, hipStream_t s);
hipError_t launch_reactors(const ReactorPool &p, int model, const Tab *tb, int max_steps, hipStream_t s);
// the same step with one wave per slot, over slots [0, n_slots); the servers take it while at most
// kReactorWaveSlots slots are in flight (one lane per slot is the faster choice beyond: a wave per slot
// costs about 18 times the SIMD time of a lane per slot, and wins only while lanes would idle)
constexpr int kReactorWaveSlots = 2048;
hipError_t launch_reactors_wave(const ReactorPool &p, int model, const Tab *tb, int max_steps, int n_slots,
                                hipStream_t s);
hipError_t launch_reactor_admit(const ReactorPool &p, int model, const Tab *tb, const double *req, int n_req,
                                hipStream_t s);
// primitive-variable path (reconstruction + LF / AUSM+ / HLLC) and Roe, one direction
hipError_t launch_inviscid_pv(const Zone &z, const Run &r, const Tab *tb, int dir, hipStream_t s);
hipError_t launch_viscous_flux(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
// store: replace dq instead of adding to it (the stage's first contribution, dq not cleared before)
hipError_t launch_viscous_derivative(const Zone &z, const Run &r, hipStream_t s, bool store = false);
// 2nd-order face scheme, all directions (flux faces in fFlux, then the difference)
hipError_t launch_viscous2(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
hipError_t launch_periodic_flux(const Zone &z, const Run &r, const corefl_boundary &b, hipStream_t s);
hipError_t launch_bc_outflow(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b, hipStream_t s);
hipError_t launch_bc_periodic(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b, hipStream_t s);
// both periodic faces of one axis in one launch (lo: the face at index 0 of the axis); the same result as the
// two faces' launches in sequence when the block holds at least 2 ngg + 1 nodes along the axis
hipError_t launch_bc_periodic_pair(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &lo,
                                  hipStream_t s);
hipError_t launch_periodic_flux_pair(const Zone &z, const Run &r, const corefl_boundary &lo, hipStream_t s);
hipError_t launch_bc_symmetry(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b, hipStream_t s);
hipError_t launch_bc_inflow(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b,
                            const corefl_bc_condition &c, hipStream_t s);
hipError_t launch_bc_wall(const Zone &z, const Run &r, const Tab *tb, const corefl_boundary &b,
                          const corefl_bc_condition &c, hipStream_t s);
hipError_t launch_k_gradient(const Zone &z, const Run &r, hipStream_t s);
// limit_flow: repairs from the pre-repair state into fFlux[0..4] (flag in gFlux[0]), then applied
hipError_t launch_limit_flow(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
// inner_communication (bv, sv, cv) and inner_exchange task 1 (f/g/hFlux) of one face
hipError_t launch_inner_comm(const Zone &z, const Zone &tz, const Run &r, const Tab *tb, const InnerDev &f,
                             hipStream_t s);
hipError_t launch_inner_flux(const Zone &z, const Zone &tz, const Run &r, const InnerDev &f, hipStream_t s);
// Writes dt_local and the block minimum into *dmin (device scalar, reduced with min).
hipError_t launch_local_dt(const Zone &z, const Run &r, const Tab *tb, double *dmin, hipStream_t s);
// one boundary face of one block of a multi-block launch
struct BcItem {
  int blk;
  corefl_boundary b;
};
// multi-block launches (zones in device memory; blockIdx.y = block or face item)
hipError_t launch_zero_dq_multi(const Zone *zs, int nb, long max_n, int nv, hipStream_t s);
hipError_t launch_update_rk_multi(const Zone *zs, int nb, long max_n, const Run &r, const Tab *tb, const double *dtp,
                                  int rk, hipStream_t s, bool save_qn = false);
hipError_t launch_properties_multi(const Zone *zs, int nb, long max_padded, const Run &r, const Tab *tb,
                                   hipStream_t s);
hipError_t launch_bc_outflow_multi(const Zone *zs, const BcItem *items, int n_items, long max_face, const Run &r,
                                   const Tab *tb, hipStream_t s);
// batched 1-D shock tubes (one padded row, see TubePool)
hipError_t launch_tube_clear(int *map, const int *ranges, int n_ranges, long map_n, hipStream_t s);
hipError_t launch_tube_admit(const Zone &z, const TubePool &P, int model, const Tab *tb, const double *rec,
                             int n_rec, int *map, long map_n, hipStream_t s);
hipError_t launch_tube_cv(const Zone &z, int model, const Tab *tb, const double *rec, int n_rec, hipStream_t s);
hipError_t launch_tube_bc(const Zone &z, const TubePool &P, int model, const Tab *tb, int n_slots, hipStream_t s);
hipError_t launch_tube_dt(const Zone &z, const TubePool &P, int n_slots, hipStream_t s, double fixed_dt = 0.0);
hipError_t launch_tube_dtc(const Zone &z, const TubePool &P, const int *map, double *dtc, hipStream_t s);
hipError_t launch_tube_finish(const TubePool &P, int n_slots, hipStream_t s);
// properties over one padded row (my = mz = 1 zones with zero y/z strides)
hipError_t launch_properties_row(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
hipError_t launch_cv_from_bv_row(const Zone &z, const Run &r, const Tab *tb, hipStream_t s);
hipError_t launch_copy_planes(double *dst, const double *src, int planes, long stride, long len, hipStream_t s);
// Returns true when the species count has a compiled kernel set.
bool model_supported(int model);

}  // namespace corefl
