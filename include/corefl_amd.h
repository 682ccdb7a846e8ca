/*
 * corefl_amd.h -- C ABI of the MI355X (gfx950) implementation of COREFL's
 * per-stage right-hand side and SSP-RK3 update.
 *
 * The reference drives its hot path through C++ template free functions
 * called from RK3<mix_model>(Driver&) (src/RK.cuh:30-275).  Each entry point
 * below replaces one of those calls (or a group of them) and is what a
 * maintainer binds from the reference's Driver; the binding is shown in
 * INTEGRATION.md.
 *
 * Conventions
 *  - Return 0 on success, a negative corefl_status on failure; the message is
 *    available from corefl_last_error().  Nothing here aborts the process.
 *  - Host pointers are read during the call only and never retained.
 *  - Arrays use the reference's DZone layout: one plane per variable, i fastest,
 *    planes of (mx+2g)(my+2g)(mz+2g) doubles (src/gxl_lib/Array.cuh:242-256,
 *    309-327); dq and dt_local have no ghosts (src/Field.cu:545, 559).
 *  - One context per GPU, one host thread per context.  All work is ordered on
 *    the context's HIP stream; only downloads, corefl_sync and the CFL time step
 *    synchronise with the host.
 */
#ifndef COREFL_AMD_H
#define COREFL_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COREFL_MAX_SPEC 20
#define COREFL_SPEC_NAME_LEN 24
#define COREFL_MAX_REAC 64
#define COREFL_MAX_RANGE 4   /* NASA-9 temperature ranges per species */
#define COREFL_MAX_ELEC 16   /* electronic levels per species (two-temperature data) */

typedef enum {
  COREFL_OK = 0,
  COREFL_ERR_ARG = -1,      /* invalid argument or unsupported combination */
  COREFL_ERR_HIP = -2,      /* HIP runtime error */
  COREFL_ERR_IO = -3,       /* file missing or unreadable */
  COREFL_ERR_PARSE = -4,    /* malformed setup / Chemkin input */
  COREFL_ERR_STATE = -5     /* non-finite or non-positive state detected */
} corefl_status;

/* Chemistry / thermodynamics / transport tables; the content of the device
 * parameter snapshot DParameter (src/DParameter.cuh:13-227) built by
 * DParameter::DParameter (src/DParameter.cu:6-432) from Species/Reaction
 * (src/ChemData.h:11-115).  Species order is the mechanism's SPECIES order. */
typedef struct {
  int n_spec, n_reac;
  int nasa7or9;                                   /* 7: two-range NASA-7 */
  double mw[COREFL_MAX_SPEC], imw[COREFL_MAX_SPEC], gas_const[COREFL_MAX_SPEC];
  double t_low[COREFL_MAX_SPEC], t_mid[COREFL_MAX_SPEC], t_high[COREFL_MAX_SPEC];
  double low_coeff[COREFL_MAX_SPEC][7], high_coeff[COREFL_MAX_SPEC][7];
  int geometry[COREFL_MAX_SPEC];
  double vis_coeff[COREFL_MAX_SPEC], LJ_potent_inv[COREFL_MAX_SPEC], ZRotF298[COREFL_MAX_SPEC];
  double WjDivWi_to_One4th[COREFL_MAX_SPEC][COREFL_MAX_SPEC];
  double sqrt_WiDivWjPl1Mul8[COREFL_MAX_SPEC][COREFL_MAX_SPEC];
  double kb_over_eps_jk[COREFL_MAX_SPEC][COREFL_MAX_SPEC];
  double binary_diffusivity_coeff[COREFL_MAX_SPEC][COREFL_MAX_SPEC];
  int reac_type[COREFL_MAX_REAC], rev_type[COREFL_MAX_REAC], order[COREFL_MAX_REAC];
  int stoi_f[COREFL_MAX_REAC][COREFL_MAX_SPEC], stoi_b[COREFL_MAX_REAC][COREFL_MAX_SPEC];
  double A[COREFL_MAX_REAC], b[COREFL_MAX_REAC], Ea[COREFL_MAX_REAC];
  double A2[COREFL_MAX_REAC], b2[COREFL_MAX_REAC], Ea2[COREFL_MAX_REAC];
  double third_body[COREFL_MAX_REAC][COREFL_MAX_SPEC];
  double troe_alpha[COREFL_MAX_REAC], troe_t1[COREFL_MAX_REAC], troe_t2[COREFL_MAX_REAC],
      troe_t3[COREFL_MAX_REAC];
  /* NASA-9 (nasa7or9 == 9; src/ChemData.cpp:865-1034, DParameter.cu:119-132): species l has
   * n_range[l] ranges t_cut[l][j] .. t_cut[l][j+1], each with a1..a7, b1, b2 in poly9[l][j] */
  int n_range[COREFL_MAX_SPEC];
  double t_cut[COREFL_MAX_SPEC][COREFL_MAX_RANGE + 1];
  double poly9[COREFL_MAX_SPEC][COREFL_MAX_RANGE][9];
  /* Two-temperature data (the two_temperature file, src/ChemData.cpp:130-232); two_temperature
   * is 1 when it was loaded.  theta_v: characteristic vibrational temperature (0: none);
   * elec_theta/elec_g: electronic levels; lt_a/lt_b: Millikan-White pair coefficients
   * (0: the default correlation); park_sigma: Park's limiting cross-section. */
  int two_temperature;
  double theta_v[COREFL_MAX_SPEC];
  int n_elec[COREFL_MAX_SPEC];
  double elec_theta[COREFL_MAX_SPEC][COREFL_MAX_ELEC], elec_g[COREFL_MAX_SPEC][COREFL_MAX_ELEC];
  double lt_a[COREFL_MAX_SPEC][COREFL_MAX_SPEC], lt_b[COREFL_MAX_SPEC][COREFL_MAX_SPEC];
  double park_sigma[COREFL_MAX_SPEC];
  /* Reaction (control) temperature T^a Tve^b, smoothed, per reaction for the forward (tcf) and
   * backward (tcb) rates when trx_model is 1 (src/FiniteRateChem.cu:1490-1503, ChemData.cpp:1248-1262) */
  int trx_model;
  double tcf_a[COREFL_MAX_REAC], tcf_b[COREFL_MAX_REAC], tcb_a[COREFL_MAX_REAC], tcb_b[COREFL_MAX_REAC];
  /* species names as the mechanism's SPECIES section lists them, upper case (Species::spec_list,
   * src/ChemData.cpp:11-60); they name the scalars in solution files */
  char species_name[COREFL_MAX_SPEC][COREFL_SPEC_NAME_LEN];
} corefl_chem_tables;

/* Run switches: the scalars of DParameter that the path reads
 * (src/DParameter.cu:6-60) plus the step controls read by RK3
 * (src/RK.cuh:66-80).  Derived counts follow Parameter::deduce_sim_info
 * (src/Parameter.cpp:571-687). */
typedef struct {
  int mixture;               /* 0: Air perfect gas (gamma 1.4), 1: species mixture */
  int n_spec, n_var, n_scalar, n_scalar_transported;
  int inviscid_scheme;       /* 52/72: characteristic, 51/71 component-wise WENO5/7; 1 LF, 2 Roe, 3 AUSM+, 4 HLLC */
  int viscous_order;         /* 0: inviscid, 8: 8th-order collocated */
  int reaction;              /* 0: frozen, 1: finite-rate chemistry (explicit) */
  int gradp_in_diffusion;    /* gradPInDiffusionFlux */
  int fixed_time_step;       /* 1: dt given, 0: dt = CFL-limited global minimum */
  double dt, cfl, prandtl;
  double rho_ref, v_ref;     /* reference state; WENO eps = 1e-40 (rho_ref v_ref)^2 */
  /* primitive-variable path (inviscid_scheme 1 LF (Air), 2 Roe, 3 AUSM+, 4 HLLC),
   * src/InviscidScheme.cu:28-787: reconstruction 1 first order, 2 MUSCL,
   * 3 NND2; limiter 0 or 1 (minmod); Roe entropy-fix factor (default 0.125) */
  int reconstruction, limiter;
  double entropy_fix_factor;
  /* limit_flow after the update (src/RK.cuh:208-209, src/TimeAdvanceFunc.cuh:129-232):
   * bounds rho in [1e-6, 1e3] rho_ref, u,v,w within 1000 v_ref, p in [1e-6, 1e3] p_ref */
  int limit_flow;
  double p_ref;
  /* finite-rate chemistry kernel: 0 compiles one for the mechanism when the
   * context is created (constants inlined, reactions unrolled; hipRTC), 1 uses
   * the table-driven kernel.  Both evaluate src/FiniteRateChem.cu:1406-1681. */
  int chem_kernel;
  /* two-temperature model (the COREFL_ENABLE_TWO_TEMPERATURE build with a two_temperature_file):
   * scalar n_spec is the vibrational-electronic energy E_ve, n_var = 6 + n_spec; vt_relaxation is
   * the setup's enable_vt_relaxation (src/RK.cuh:200-203) */
  int two_temperature;
  int vt_relaxation;
  /* WENO options (src/WENO.cu): positive_preserving limits the species (and E_ve) fluxes of the
   * component-wise kernels and of the x characteristic kernel; a shockSensor_threshold above 1e-10
   * turns the WENO schemes into the hybrid UD/WENO schemes (src/Parameter.cpp:252-270), with the
   * sensor (shock_sensor 0 modified Ducros with shockSensor_epsilon, 1 modified Jameson, 2 density
   * and pressure jump; src/FieldOperation.cu:122-228) computed once per step */
  int positive_preserving;
  int shock_sensor;
  double sensor_eps, sensor_threshold;
  /* 1: corefl_step records the whole RK3 step once as a HIP graph (blocks' work on concurrent streams,
   * dt kept in device memory) and replays it; rebuilt when blocks are added or a fixed dt changes */
  int step_graph;
  /* 1: corefl_initialize times the compiled tile shapes of the characteristic WENO sweeps on each
   * block (its own extents) and keeps the fastest per block and direction; every shape gives the
   * same fluxes bit for bit.  0: the default shapes. */
  int autotune;
  /* temporal_scheme: 3 SSP-RK3 (default), 4 Wu's operator splitting (src/StrangSplitting.cu:687-865: the
   * transport part of dq frozen as Cn, a reaction sub-step over dt with chemistry sub-steps from the
   * chemical time scale, then an RK3 sub-step over dt / 2 of dq - Cn; a one-temperature reacting
   * mixture).  chem_advance_method: its reaction sub-step, 0 explicit, 1 point-implicit with the full
   * Jacobian, 3 SSP-RK3 (src/StrangSplitting.cu:43-228).  0 in either means the default. */
  int temporal_scheme;
  int chem_advance_method;
} corefl_run_desc;

/* One boundary face of a block (src/Mesh.h Boundary, src/Mesh.cpp:14-32):
 * face 0/1/2 = i/j/k, direction -1 (min face) or +1 (max face), interior
 * index ranges inclusive, type = the boundary file's label (see
 * corefl_bc_condition below). */
typedef struct {
  int face, direction, type;
  int range_start[3], range_end[3];
} corefl_boundary;

/* The condition a boundary label stands for: the setup's struct for that
 * label (src/BoundCond.h Wall 90-140, Inflow 22-88; built by
 * src/BoundCond.cpp:49-600).  Register one with corefl_bc_set before adding
 * blocks that use the label.  Without a registered condition, label 2 is an
 * isothermal wall at 300 K (the reference's default), 3 symmetry, 6 outflow
 * and 8 periodic. */
#define COREFL_BC_WALL 2
#define COREFL_BC_SYMMETRY 3
#define COREFL_BC_INFLOW 5
#define COREFL_BC_OUTFLOW 6
#define COREFL_BC_PERIODIC 8
typedef struct {
  int label;
  int kind;               /* COREFL_BC_* */
  int thermal_type;       /* wall: 0 isothermal, 1 adiabatic (Wall::ThermalType order) */
  int inflow_sub;         /* inflow: 1 subsonic characteristic treatment (Inflow::inflow_sub) */
  double wall_temperature;
  double density, u, v, w, pressure, temperature; /* constant inflow state (inflow_type 0) */
  double mw;              /* inflow mixture molecular weight (mixture runs) */
  double sv[COREFL_MAX_SPEC]; /* inflow scalars (mass fractions; with two temperatures E_ve last, which
                               corefl_bc_set derives from temperature_ve as Inflow's constructor does) */
  double temperature_ve;  /* two temperatures: inflow / isothermal-wall Tve ("Tve" key; <= 0: the temperature) */
} corefl_bc_condition;

/* One 1-to-1 interface face, the two lines of boundary_condition/inner%4d.txt
 * (read by Mesh::read_inner_interface, src/Mesh.cpp:927-958; registered by
 * InnerFace::register_boundary, src/Mesh.cpp:39-172): the source and target
 * ranges x1 x2 y1 y2 z1 z2, 1-based, the normal axis repeated (1: low face),
 * the corresponding tangential axes negated (3-D), running direction by the
 * order of start and end; target_block is the target line's 1-based block
 * number (blocks are numbered in the order they are added). */
typedef struct {
  int src[6];
  int tar[6];
  int target_block;
} corefl_inner_face;

/* Block geometry (src/Mesh.h Block): metric (9 planes: xi_x..zeta_z) and
 * jacobian over the padded extent, boundary faces in the order the reference
 * applies them, inner faces in file order.  bType masks (src/Mesh.cpp:856-916)
 * are derived from the faces: periodic and inner faces count as interior. */
typedef struct {
  int mx, my, mz, ngg;
  const double *metric;
  const double *jac;
  int n_boundary;
  const corefl_boundary *boundary;
  int n_inner;
  const corefl_inner_face *inner;
} corefl_block_desc;

/* Fields reachable through corefl_field_download / corefl_field_upload. */
typedef enum {
  COREFL_F_CV = 0, COREFL_F_BV, COREFL_F_SV, COREFL_F_QN, COREFL_F_DQ, COREFL_F_FFLUX, COREFL_F_GFLUX,
  COREFL_F_HFLUX, COREFL_F_METRIC, COREFL_F_JAC, COREFL_F_MUL, COREFL_F_KAPPA, COREFL_F_GAMMA,
  COREFL_F_ACOUSTIC, COREFL_F_MACH, COREFL_F_RHOD, COREFL_F_RTS, COREFL_F_DT_LOCAL,
  COREFL_F_TVE,      /* two-temperature: temperature_ve (src/Field.h) */
  COREFL_F_KAPPA_VE, /* two-temperature: thermal_conductivity_ve */
  COREFL_F_SHOCK_SENSOR, /* shock_sensor (hybrid WENO only) */
  COREFL_F_COUNT
} corefl_field;

/* Per-stage operators; the parity unit of test. */
typedef enum {
  COREFL_OP_ZERO_DQ = 0,     /* cudaMemset(dq), src/RK.cuh:166-169 */
  COREFL_OP_VISCOUS,         /* compute_viscous_flux, src/ViscousScheme.cu:24-84 */
  COREFL_OP_CHEMISTRY,       /* finite_rate_chemistry, src/FiniteRateChem.cu:1406-1488 */
  COREFL_OP_INVISCID_X,      /* compute_convective_term_weno_x + derivative_x, src/WENO.cu:831-1559 */
  COREFL_OP_INVISCID_Y,      /* ..._y, src/WENO.cu:1561-2223 */
  COREFL_OP_INVISCID_Z,      /* ..._z, src/WENO.cu:2225-2886 */
  COREFL_OP_UPDATE_RK,       /* update_cv_and_bv_rk, src/RK.cuh:277-367 */
  COREFL_OP_BC,              /* DBoundCond::apply_boundary_conditions, src/BoundCond.cu:3846-3997 */
  COREFL_OP_K_GRADIENT,      /* eliminate_k_gradient (2-D), src/FieldOperation.cu:96-121 */
  COREFL_OP_PROPERTIES,      /* update_physical_properties, src/FieldOperation.cuh:183-247 */
  COREFL_OP_CV_FROM_BV,      /* compute_cv_from_bv, src/FieldOperation.cuh:136-161 */
  COREFL_OP_LOCAL_DT,        /* local_time_step + global min, src/TimeAdvanceFunc.cuh:28-127 */
  COREFL_OP_COPY_QN,         /* qn <- cv, src/RK.cuh:126-129 */
  COREFL_OP_HALO,            /* data_communication (inner faces, all blocks), src/DataCommunication.cu:6-124 */
  COREFL_OP_LIMIT_FLOW,      /* limit_flow, src/TimeAdvanceFunc.cuh:129-232 */
  COREFL_OP_SHOCK_SENSOR,    /* compute_shock_sensor, src/FieldOperation.cu:207-228 */
  COREFL_OP_WU_CN,           /* Cn <- dq (compute_cn's copy, src/StrangSplitting.cu:36-39) */
  COREFL_OP_WU_REACTION,     /* wu_reaction_step over dt, src/StrangSplitting.cu:377-425 */
  COREFL_OP_UPDATE_RK_WU,    /* update_cv_and_bv_rk_wu (rk, dt), src/StrangSplitting.cu:508-566 */
  COREFL_OP_COUNT
} corefl_op_id;

typedef struct corefl_ctx corefl_ctx;

/* Chemkin readers (src/ChemData.cpp:11-1493): mechanism, NASA-7 therm.dat and
 * transport file into tables.  Replaces Species(Parameter&)/Reaction(...). */
int corefl_chem_load(const char *mechanism_path, const char *therm_path, const char *transport_path,
                     corefl_chem_tables *out);

/* The same with the setup keys therm_nasa_7_9 (7 or 9) and two_temperature_file (NULL: none),
 * as Species(Parameter&) reads them (src/ChemData.cpp:84-121, 130-232). */
int corefl_chem_load_ex(const char *mechanism_path, const char *therm_path, const char *transport_path,
                        int nasa_7_or_9, const char *two_temperature_path, corefl_chem_tables *out);

/* Generates and compiles (hipRTC, gfx950) the chemistry kernel that a context
 * with chem_kernel 0 builds for these tables, without loading it: no device is
 * needed.  Returns 0, or COREFL_ERR_HIP with the compiler log in log.  No
 * counterpart in the reference, whose chemistry kernel reads the mechanism from
 * DParameter (src/FiniteRateChem.cu:1406-1681). */
int corefl_chem_kernel_check(const corefl_chem_tables *tables, char *log, size_t log_cap);

/* Reads a reference-format setup.txt (src/Parameter.cpp:38-184) and fills the
 * run switches for the keys the path uses (defaults src/Parameter.cpp:358-546).
 * When the setup names a mechanism the tables are loaded from case_dir/input. */
int corefl_setup_load(const char *case_dir, corefl_run_desc *run, corefl_chem_tables *tables);

/* Context: replaces the Driver's DParameter upload (src/Driver.cu:97-99). */
int corefl_ctx_create(const corefl_run_desc *run, const corefl_chem_tables *tables, corefl_ctx **out);
void corefl_ctx_destroy(corefl_ctx *ctx);
const char *corefl_last_error(const corefl_ctx *ctx);

/* Registers the condition of a boundary label (replaces DBoundCond's
 * per-label tables, src/BoundCond.cpp). */
int corefl_bc_set(corefl_ctx *ctx, const corefl_bc_condition *cond);

/* Block storage: replaces Field::setup_device_memory (src/Field.cu:439-681). */
int corefl_block_add(corefl_ctx *ctx, const corefl_block_desc *desc, int *block_id);
/* Upload basic variables bv[6] and scalars sv[n_scalar] over the padded extent
 * (initialize_basic_variables output, src/Initialize.cu:74-160). */
int corefl_state_upload(corefl_ctx *ctx, int blk, const double *bv, const double *sv);
int corefl_field_download(corefl_ctx *ctx, int blk, int field, double *out);
int corefl_field_upload(corefl_ctx *ctx, int blk, int field, const double *in);
/* Number of doubles corefl_field_download writes for `field`. */
long corefl_field_size(const corefl_ctx *ctx, int blk, int field);

/* Driver::initialize_computation (src/Driver.cu:126-211) followed by the RK3
 * preamble compute_cv_from_bv (src/RK.cuh:47-58). */
int corefl_initialize(corefl_ctx *ctx);
/* One operator on one block (blk < 0: all blocks).  COREFL_OP_VISCOUS and
 * COREFL_OP_HALO couple the blocks and always run on all of them. */
int corefl_op(corefl_ctx *ctx, int op, int blk, int rk, double dt);
/* One RK3 stage on all blocks, src/RK.cuh:164-236.  The update kernel flags
 * any interior point whose density, pressure or temperature is non-finite or
 * not positive, or whose mass fractions are non-finite (the states at which
 * the reference's residual check stops the run, src/Residual.cu:97-118,
 * 211-243); corefl_stage, corefl_step and corefl_sync then return
 * COREFL_ERR_STATE (without running further work) until corefl_initialize or
 * corefl_state_upload clears the flag. */
int corefl_stage(corefl_ctx *ctx, int rk, double dt);
/* Tile shape (0 or 1) each direction's characteristic WENO sweep uses on block blk; the setter pins
 * shapes (e.g. found offline), which a later corefl_initialize with autotune re-times. */
int corefl_block_tiles(const corefl_ctx *ctx, int blk, int shapes[3]);
int corefl_block_set_tiles(corefl_ctx *ctx, int blk, const int shapes[3]);
/* One RK3 step: qn <- cv, dt (fixed or CFL), three stages, src/RK.cuh:107-272.
 * dt_inout: in = fixed dt (ignored when the run is CFL-based), out = dt used. */
int corefl_step(corefl_ctx *ctx, double *dt_inout);
/* Waits for the context's stream; COREFL_ERR_STATE when a state was flagged. */
int corefl_sync(corefl_ctx *ctx);

/* Live per-operator timing with HIP events on the context stream (the
 * reference's profile_rk_breakdown, src/RK.cuh:90-101, without host syncs). */
int corefl_profile_begin(corefl_ctx *ctx);
/* Fills ms[COREFL_OP_COUNT] (summed milliseconds) and calls[COREFL_OP_COUNT]. */
int corefl_profile_end(corefl_ctx *ctx, double *ms, long *calls);

/* ---- Solution files (Tecplot binary V112, all values FP64) ----
 * corefl_plt_write replaces FieldIO<>'s output/flowfield.plt (write_header, write_common_data_section,
 * src/FieldIO.h:96-420, and print_field 709-800): one ordered zone per block, "zone <b>", with
 * (mx + 2 ngg_out)(my + 2 ngg_out)(mz + 2 ngg_out) nodes (0 <= ngg_out <= ngg), variables x, y, z,
 * density, u, v, w, pressure, temperature, the scalars named after the tables' species (then Eve with
 * two temperatures), mach (then Tve), each preceded in the data section by its min / max and written
 * i fastest.  coords[b] holds block b's node coordinates x, y, z over its padded extent (3 planes, the
 * layout of the other padded fields), or coords / coords[b] NULL for zeros.  When step_path is not NULL
 * the step count goes there as text (output/message/step.txt). */
int corefl_plt_write(corefl_ctx *ctx, const char *path, const double *const *coords, int ngg_out,
                     double solution_time, const char *step_path, int step);
/* corefl_plt_read replaces read_flowfield (src/Initialize.cu:167-420) for files whose zones have the
 * blocks' node counts without ghosts: in block order, or matched one to one by node counts (the
 * reference's same_mesh / same_block cases; otherwise COREFL_ERR_ARG, where the reference would
 * interpolate).  Variables are matched by name, case-insensitively: density / rho, u, v, w, pressure / p,
 * temperature / t, the species by the tables' names, Eve and Tve with two temperatures; others are
 * skipped.  Every bv plane and every scalar must be present.  The file's values replace the blocks'
 * interior bv, sv (and Tve); ghost values stay.  *solution_time gets the zones' solution time.  Call
 * corefl_initialize afterwards, as the reference initialises cv and the properties from bv. */
int corefl_plt_read(corefl_ctx *ctx, const char *path, double *solution_time);

/* ---- Reference case directories ----
 * The files the readGrid preprocessor and the setup leave under
 * case_dir/input/, read as the reference's Mesh does: grid/grid%4d.dat
 * (gridIsBinary 1: int block count, int mx my mz per block, then x, y, z of
 * every block, i fastest) or grid%4d.grd (the same as text), scaled by
 * gridScale (Mesh::read_grid, src/Mesh.cpp:692-835);
 * boundary_condition/boundary%4d.txt (per block a count, then
 * "i1 i2 j1 j2 k1 k2 label" lines, 0-based; src/Mesh.cpp:837-925) and
 * inner%4d.txt (src/Mesh.cpp:927-962), %4d being the rank.  The nodes are
 * extended into ngg + 1 ghost layers (mirrored across faces, copied from the
 * neighbour across interfaces, parallelogram rule at edges and corners;
 * Mesh::init_ghost_grid, src/Mesh.cpp:1013-1217) and the Jacobian and metrics
 * taken by central differences over the padded extent
 * (Block::compute_jac_metric, src/Mesh.cpp:244-528).  ngg follows the
 * setup's schemes (src/Parameter.cpp:216-321). */
typedef struct corefl_mesh corefl_mesh;
int corefl_mesh_load(const char *case_dir, int rank, corefl_mesh **out);
int corefl_mesh_n_block(const corefl_mesh *mesh);
/* The block's geometry for corefl_block_add; its pointers stay valid while the mesh lives. */
int corefl_mesh_block(const corefl_mesh *mesh, int blk, corefl_block_desc *desc);
/* Node coordinates over the padded extent: three planes x, y, z in the DZone layout. */
int corefl_mesh_coords(const corefl_mesh *mesh, int blk, double *xyz);
/* initialize_basic_variables (src/Initialize.cu:118-160, src/Field.cu:289-360): bv[6], sv[n_spec] and
 * (two temperatures, may be NULL) Tve over the padded extent, each node from the first init_cond_<n> box
 * that holds it or else default_init, every state derived from its setup struct as Inflow's constructor
 * does (src/BoundCond.cpp:150-330). */
int corefl_mesh_initial_state(corefl_mesh *mesh, int blk, const corefl_chem_tables *tables, double *bv,
                              double *sv, double *tve);
/* The conditions of the setup's boundary_conditions structs (wall, inflow, outflow, symmetry, periodic)
 * for corefl_bc_set; returns how many there are (conds receives up to max). */
int corefl_mesh_conditions(corefl_mesh *mesh, const corefl_chem_tables *tables, corefl_bc_condition *conds,
                           int max);
const char *corefl_mesh_error(const corefl_mesh *mesh);
void corefl_mesh_free(corefl_mesh *mesh);

/* ---- Request server for 0-D constant-volume reactors ----
 * No counterpart in the reference (whose const_volume_reactor,
 * src/ConstVolumeReactor.cpp:320-775, integrates one case per run): reactors
 * arrive as requests, are admitted into device-resident slots while the slot
 * budget allows, advance together (one launch per progress call, every active
 * slot stepped with the reference's schemes) and leave as results; requests
 * that arrive while others are in flight join at the next progress call. */
typedef struct corefl_server corefl_server;

typedef struct {
  double temperature, pressure;  /* initial T (K) and p (Pa) */
  double y[COREFL_MAX_SPEC];     /* initial mass fractions */
  double t_end;                  /* integrate until time reaches t_end (s) */
  double dt;                     /* fixed step (schemes 0, 3), first step (scheme 1) */
  int scheme;                    /* 0 point-implicit with the diagonal Jacobian, 1 ROK4E, 3 SSP-RK3 */
  int adaptive;                  /* ROK4E: error-controlled steps (RTol / ATol) */
  double rtol, atol;
  int conserve_mass;             /* scheme 0: renormalise the mass fractions after the update */
  long max_steps;                /* > 0: a request whose accepted + rejected steps reach it before t_end fails */
} corefl_reactor_request;

/* Why a served request failed (result status COREFL_ERR_STATE). */
#define COREFL_FAIL_NONFINITE 1  /* temperature, time, step or a partial density not finite */
#define COREFL_FAIL_STEP 2       /* step size not positive, or too small to advance the time */
#define COREFL_FAIL_CAP 3        /* max_steps reached before t_end */

typedef struct {
  double temperature, pressure, time;
  double y[COREFL_MAX_SPEC];
  long steps, rejected;          /* accepted and rejected steps */
  int status;                    /* 0: reached t_end; COREFL_ERR_STATE: failed, state as last stored */
  int reason;                    /* COREFL_FAIL_* when failed */
} corefl_reactor_result;

/* budget_bytes of device memory fixes the number of slots; steps_per_launch
 * bounds the steps one progress call takes per slot. */
int corefl_serve_open(const corefl_chem_tables *tables, size_t budget_bytes, int steps_per_launch,
                      corefl_server **out);
/* Queues a request; returns its ticket (>= 0) or a negative status. */
long corefl_serve_submit(corefl_server *srv, const corefl_reactor_request *req);
/* Admits queued requests into free slots, advances every active slot by one
 * launch and collects the finished ones; *pending = requests queued or in flight. */
int corefl_serve_progress(corefl_server *srv, int *pending);
/* 1: finished or failed (res filled, res->status tells which; ticket released),
 * 0: queued or in flight, < 0: unknown ticket.  A failed request frees its slot
 * and leaves the other requests untouched. */
int corefl_serve_poll(corefl_server *srv, long ticket, corefl_reactor_result *res);
/* Moves up to max finished results out (tickets[i], res[i]); returns how many. */
int corefl_serve_drain(corefl_server *srv, long *tickets, corefl_reactor_result *res, int max);
int corefl_serve_capacity(const corefl_server *srv);
const char *corefl_serve_error(const corefl_server *srv);
void corefl_serve_close(corefl_server *srv);

/* Allocation-free batch of reactors.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture of its own work, tight serving loops, harnesses that forbid
 * hipMalloc).  It runs n requests (0 <= n <= COREFL_NOALLOC_MAX_REACTORS) to
 * t_end or failure with the reactor server's kernels, admission and result
 * rules, steps_per_launch steps per launch, and fills res[0..n) exactly as
 * corefl_serve_poll would: a request's result is bit-identical to the
 * server's.  Device working memory is one fixed array per GPU (3 MiB, held
 * until the process ends), tables and slots included; host staging is static
 * too.  Nothing is allocated: no hipMalloc, hipFree or allocating library call.
 * Runs on the current device and the null stream and returns when done.  Every
 * call in a process shares the array of its GPU and the host staging, which
 * keep the last call's contents (each call rewrites all it reads): calls that
 * overlap, from two host threads, race; serialise them.  Returns 0 or a
 * negative status (COREFL_ERR_ARG for bad sizes or requests, COREFL_ERR_HIP);
 * it keeps no error message. */
#define COREFL_NOALLOC_MAX_REACTORS 4096
int corefl_serve_run_noalloc(const corefl_chem_tables *tables, const corefl_reactor_request *req, int n,
                             int steps_per_launch, corefl_reactor_result *res);

/* ---- Serving 1-D reactive shock tubes (no counterpart in the reference,
 * which advances one grid per process; each tube is what the reference runs
 * as a one-row block: the same inviscid scheme, chemistry, RK3 stage order,
 * outflow / symmetry boundary kernels and CFL step from local_time_step,
 * src/TimeAdvanceFunc.cuh:28-127).  Tubes of different sizes lie end to end
 * in one device-resident row sized by the budget; a progress call admits
 * queued tubes into free stretches of the row and advances every tube in
 * flight by a number of RK3 steps, each tube on its own CFL step and its own
 * clock, with one launch per operator for all of them.  A tube's result is
 * the same alone and among others. ---- */
typedef struct corefl_tube_server corefl_tube_server;

typedef struct {
  int n_cells;                   /* grid nodes along the tube (>= 16) */
  int n_left;                    /* nodes [0, n_left) start in state 0, the rest in state 1 */
  double length;                 /* m; node spacing length / (n_cells - 1), also in y and z */
  double t_end;                  /* s */
  int bc_left, bc_right;         /* COREFL_BC_OUTFLOW or COREFL_BC_SYMMETRY (reflecting end) */
  double temperature[2], pressure[2], velocity[2];  /* K, Pa, m/s; [0] left, [1] right */
  double y[2][COREFL_MAX_SPEC];  /* mass fractions of the two states */
  long max_steps;                /* > 0: a tube still short of t_end after this many steps fails */
} corefl_tube_request;

typedef struct {
  double time;                   /* reached time (t_end) */
  long steps;                    /* RK3 steps taken */
  int n_cells;
  int status;                    /* 0: reached t_end; COREFL_ERR_STATE: failed */
  int reason;                    /* COREFL_FAIL_NONFINITE (a node's local step not finite and positive) or
                                    COREFL_FAIL_CAP */
} corefl_tube_result;

/* run: inviscid_scheme 51/52/71/72 (ngg 3 for WENO5, 4 for WENO7), cfl,
 * reaction, chem_kernel, rho_ref / v_ref for the WENO epsilon; a one-
 * temperature species mixture without viscous terms, sensor or limiter.
 * budget_bytes of device memory fixes the row's length; steps_per_progress
 * is the number of RK3 steps one progress call takes. */
int corefl_tubes_open(const corefl_run_desc *run, const corefl_chem_tables *tables, size_t budget_bytes,
                      int steps_per_progress, corefl_tube_server **out);
long corefl_tubes_submit(corefl_tube_server *srv, const corefl_tube_request *req);
int corefl_tubes_progress(corefl_tube_server *srv, int *pending);
/* 1: finished or failed (res filled, res->status tells which; profile, when not null, receives (6 + n_spec) rows
 * of n_cells values: rho, u, v, w, p, T, Y_0 .. Y_{ns-1}; ticket released),
 * 0: queued or in flight, < 0: unknown ticket or a profile buffer too small. */
int corefl_tubes_poll(corefl_tube_server *srv, long ticket, corefl_tube_result *res, double *profile,
                      long profile_len);
/* Lists up to max finished tickets not yet polled (in no particular order); returns how many. */
int corefl_tubes_ready(const corefl_tube_server *srv, long *tickets, int max);
long corefl_tubes_capacity(const corefl_tube_server *srv); /* nodes the row holds (tube + 2 ngg each) */
const char *corefl_tubes_error(const corefl_tube_server *srv);
void corefl_tubes_close(corefl_tube_server *srv);

#ifdef __cplusplus
}
#endif
#endif
