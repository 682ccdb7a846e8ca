/*
 * corefl_oracle.h -- CPU restatement of COREFL's per-stage RHS and SSP-RK3 update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product never links, calls or falls back to it.
 *
 * Every routine cites the reference function it restates (paths relative to the
 * reference repository root).  Data layout follows the reference's DZone:
 * one plane per variable, i fastest, planes of size (mx+2g)(my+2g)(mz+2g)
 * (src/gxl_lib/Array.cuh:242-256, 309-327); dq and dt_local carry no ghosts
 * (src/Field.cu:545, 559).
 */
#ifndef COREFL_ORACLE_H
#define COREFL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OC_MAX_SPEC 20
#define OC_MAX_REAC 64
#define OC_MAX_RANGE 4
#define OC_MAX_ELEC 16

/* Chemistry / thermo / transport tables (src/DParameter.cuh:13-227). */
typedef struct {
  int n_spec, n_reac;
  int nasa7or9; /* 7: two-range NASA-7 (Combustion2Part build); 9: multi-range NASA-9 */
  double mw[OC_MAX_SPEC], imw[OC_MAX_SPEC], gas_const[OC_MAX_SPEC];
  double t_low[OC_MAX_SPEC], t_mid[OC_MAX_SPEC], t_high[OC_MAX_SPEC];
  double low_coeff[OC_MAX_SPEC][7], high_coeff[OC_MAX_SPEC][7];
  int geometry[OC_MAX_SPEC];
  double vis_coeff[OC_MAX_SPEC], LJ_potent_inv[OC_MAX_SPEC], ZRotF298[OC_MAX_SPEC];
  double WjDivWi_to_One4th[OC_MAX_SPEC][OC_MAX_SPEC];
  double sqrt_WiDivWjPl1Mul8[OC_MAX_SPEC][OC_MAX_SPEC];
  double kb_over_eps_jk[OC_MAX_SPEC][OC_MAX_SPEC];
  double binary_diffusivity_coeff[OC_MAX_SPEC][OC_MAX_SPEC];
  int reac_type[OC_MAX_REAC], rev_type[OC_MAX_REAC], order[OC_MAX_REAC];
  int stoi_f[OC_MAX_REAC][OC_MAX_SPEC], stoi_b[OC_MAX_REAC][OC_MAX_SPEC];
  double A[OC_MAX_REAC], b[OC_MAX_REAC], Ea[OC_MAX_REAC];
  double A2[OC_MAX_REAC], b2[OC_MAX_REAC], Ea2[OC_MAX_REAC];
  double third_body[OC_MAX_REAC][OC_MAX_SPEC];
  double troe_alpha[OC_MAX_REAC], troe_t1[OC_MAX_REAC], troe_t2[OC_MAX_REAC], troe_t3[OC_MAX_REAC];
  /* NASA-9 ranges and coefficients a1..a7, b1, b2 (src/DParameter.cu:119-132) */
  int n_range[OC_MAX_SPEC];
  double t_cut[OC_MAX_SPEC][OC_MAX_RANGE + 1];
  double poly9[OC_MAX_SPEC][OC_MAX_RANGE][9];
  /* two-temperature data (src/DParameter.cuh two-temperature members, ChemData.cpp:130-232) */
  int two_temperature;
  double theta_v[OC_MAX_SPEC];
  int n_elec[OC_MAX_SPEC];
  double elec_theta[OC_MAX_SPEC][OC_MAX_ELEC], elec_g[OC_MAX_SPEC][OC_MAX_ELEC];
  double lt_a[OC_MAX_SPEC][OC_MAX_SPEC], lt_b[OC_MAX_SPEC][OC_MAX_SPEC];
  double park_sigma[OC_MAX_SPEC];
  int trx_model; /* two_temperature_reaction_temperature */
  double tcf_a[OC_MAX_REAC], tcf_b[OC_MAX_REAC], tcb_a[OC_MAX_REAC], tcb_b[OC_MAX_REAC];
} oc_tables;

/* Run-time switches (subset of DParameter, src/DParameter.cu:6-60). */
typedef struct {
  int mixture;             /* 0: Air perfect gas, 1: species mixture */
  int n_spec, n_var, n_scalar, n_scalar_transported;
  int inviscid_scheme;     /* 52 / 72 characteristic, 51 / 71 component-wise WENO5 / WENO7; 1 LF, 2 Roe, 3 AUSM+, 4 HLLC */
  int viscous_order;       /* 0, 2 or 8 */
  int reaction;            /* 0 / 1 */
  int gradp_in_diffusion;
  double prandtl;
  double cfl;
  double weno_eps_scale;   /* (rho_ref * v_ref)^2, src/DParameter.cu:29-31 */
  int reconstruction;      /* primitive-variable path: 1 first order, 2 MUSCL, 3 NND2 */
  int limiter;             /* 0 / 1 (both minmod, src/Limiter.cuh:14-16) */
  double entropy_fix_factor; /* Roe path, default 0.125 (src/Parameter.cpp:396) */
  int limit_flow;          /* repair unphysical points after the update (src/RK.cuh:208-209) */
  double rho_inf, v_inf, p_inf; /* limit_flow bounds, src/DParameter.cu:371-387 */
  int two_temperature;     /* i_eve = n_spec >= 0: E_ve is scalar n_spec (COREFL_ENABLE_TWO_TEMPERATURE) */
  int vt_relaxation;       /* enable_vt_relaxation */
  int positive_preserving; /* species positivity limiter of the WENO fluxes (src/WENO.cu:991, 1459, 1719, 2384) */
  double pp_dt;            /* DParameter::dt: the fixed time step, 0 otherwise (then dt_local is used) */
  int shock_sensor;        /* 0 modified Ducros, 1 modified Jameson, 2 density/pressure jump */
  double sensor_eps;       /* shockSensor_epsilon */
  double sensor_threshold; /* shockSensor_threshold: > 1e-10 selects the hybrid linear/WENO weights */
  int chem_advance_method; /* Wu splitting's reaction sub-step: 0 explicit, 1 point-implicit (full Jacobian),
                              3 SSP-RK3 (src/StrangSplitting.cu:43-228) */
} oc_run;

/* One boundary face (src/Mesh.h Boundary; src/Mesh.cpp:14-32) with the
 * condition its label names in the setup (src/BoundCond.h Wall/Inflow). */
typedef struct {
  int face, direction, type; /* type: 2 wall, 3 symmetry, 5 inflow, 6 outflow, 8 periodic */
  int range_start[3], range_end[3];
  int thermal_type;          /* wall: 0 isothermal, 1 adiabatic (Wall::ThermalType order) */
  int inflow_sub;            /* inflow: 1 subsonic characteristic treatment */
  double wall_temperature;
  double density, u, v, w, pressure, temperature, mw; /* constant inflow state */
  double sv[OC_MAX_SPEC];
  double temperature_ve;     /* two temperatures: inflow / isothermal-wall Tve (<= 0: the temperature) */
} oc_boundary;

/* A 1-to-1 interface face after InnerFace::register_boundary
 * (src/Mesh.h:33-59, src/Mesh.cpp:39-172): loop ranges over the face node and
 * its ghost layers, the index map into the target block, 0-based target. */
typedef struct {
  int range_start[3], range_end[3];
  int face, direction;
  int target_start[3];
  int target_face, target_direction;
  int src_tar[3], loop_dir[3], target_loop_dir[3], n_point[3];
  int target_block;
} oc_inner_face;

/* Block storage.  All pointers are caller-owned. */
typedef struct {
  int mx, my, mz, ngg;
  double *cv, *qn, *bv, *sv, *dq;
  double *ff, *gf, *hf;
  double *metric, *jac;
  double *mul, *kappa, *gamma, *acoustic, *mach, *rhoD, *rts, *dt_local;
  double *tve, *kappa_ve; /* two-temperature: temperature_ve, thermal_conductivity_ve (padded) */
  double *shock_sensor;   /* padded; ghosts stay 0 */
  int *btype_il, *btype_ir, *btype_jl, *btype_jr, *btype_kl, *btype_kr;
  int n_boundary;
  oc_boundary *boundary;
  int n_inner;
  oc_inner_face *inner;
  double *cn;             /* Wu splitting: the transport part of dq frozen over a step (Cn, no ghosts) */
} oc_zone;

/* Registers one face of boundary_condition/inner%4d.txt: src and tar are the
 * six signed 1-based range numbers of the source and target lines
 * (x1 x2 y1 y2 z1 z2), target_block the target line's 1-based block number. */
void oc_register_inner(const int src[6], const int tar[6], int target_block, int dim, int ngg, oc_inner_face *f);

/* ---- pointwise physics (exposed for unit tests) ---- */
void oc_enthalpy(const oc_tables *tb, double t, double *h);
void oc_cp(const oc_tables *tb, double t, double *cp);
void oc_enthalpy_and_cp(const oc_tables *tb, double t, double *h, double *cp);
void oc_gibbs_div_rt(const oc_tables *tb, double t, double *g);
double oc_ve_energy(const oc_tables *tb, int i, double t);
double oc_ve_cv(const oc_tables *tb, int i, double t);
double oc_mixture_ve_energy(const oc_tables *tb, double t, const double *y, double *cv_ve);
double oc_mixture_vib_energy(const oc_tables *tb, double t, const double *y, double *cv_v);
double oc_mixture_tr_energy(const oc_tables *tb, double t, const double *y, double *cv_tr, double *cp_tr,
                            double *r_mix);
double oc_invert_tve(const oc_tables *tb, double eve_target, const double *y, double t_init);
double oc_vt_source(const oc_tables *tb, double rho, double t, double tve, const double *y, double *eve_eq,
                    double *tau_eff);
void oc_vt_relaxation(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt_stage);
void oc_chem_source(const oc_tables *tb, double rho, double t, double tve, const double *y, double *omega,
                    double *omega_d);
void oc_transport(const oc_tables *tb, double t, double mw_total, const double *y, double *cp,
                  double *mul, double *kappa, double *rhoD);
double oc_weno5(const double *vp, const double *vm, double eps, int if_shock);
double oc_weno7(const double *vp, const double *vm, double eps, int if_shock);
/* The characteristic basis of the char. WENO face between two states
 * (src/WENO.cu:1146-1457): left_m = L, right_m = R, row-major n_var x n_var. */
void oc_char_eigen(const oc_run *r, const oc_tables *tb, const double *q_l, double p_l, const double *q_r, double p_r,
                   double kx, double ky, double kz, double *left_m, double *right_m);

/* ---- per-stage operators on a block ---- */
void oc_cv_from_bv(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_update_properties(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_zero_dq(oc_zone *z, const oc_run *r);
void oc_viscous8(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_viscous2(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_chemistry(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_inviscid_dir(oc_zone *z, const oc_run *r, const oc_tables *tb, int dir);
void oc_update_rk(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk);
void oc_boundary_conditions(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_eliminate_k_gradient(oc_zone *z, const oc_run *r);
/* compute_shock_sensor, src/FieldOperation.cu:122-228, over the interior */
void oc_shock_sensor(oc_zone *z, const oc_run *r);
/* limit_flow, src/TimeAdvanceFunc.cuh:129-232, with every repair computed from
 * the state before any repair (the reference's threads read neighbours while
 * others repair, which makes adjacent bad points order-dependent). */
void oc_limit_flow(oc_zone *z, const oc_run *r, const oc_tables *tb);
double oc_local_time_step(oc_zone *z, const oc_run *r, const oc_tables *tb);

/* ---- across blocks ---- */
/* data_communication (inner faces), src/DataCommunication.cu:6-124 */
void oc_data_communication(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb);
/* compute_viscous_flux (8th order) over all blocks, src/ViscousScheme.cu:24-84 */
void oc_viscous8_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb);

/* ---- orchestration (src/Driver.cu:126-211, src/RK.cuh:30-275) ---- */
void oc_initialize(oc_zone *z, const oc_run *r, const oc_tables *tb);
void oc_stage(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk);
double oc_step(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt_fixed, int fixed);
void oc_initialize_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb);
void oc_stage_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb, double dt, int rk);
double oc_step_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb, double dt_fixed,
                      int fixed);

/* ---- Wu's operator splitting (temporal_scheme 4, src/StrangSplitting.cu:687-865) ---- */
/* compute_cn, src/StrangSplitting.cu:15-41: dq <- viscous + inviscid terms, Cn <- dq */
void oc_compute_cn_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb);
/* wu_reaction_step over the interior, src/StrangSplitting.cu:377-425 */
void oc_wu_reaction(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt);
/* update_cv_and_bv_rk_wu, src/StrangSplitting.cu:508-566 */
void oc_update_rk_wu(oc_zone *z, const oc_run *r, const oc_tables *tb, double dt, int rk);
/* One Wu step (src/StrangSplitting.cu:748-761): Cn, the reaction sub-step with
 * its boundary conditions, exchange and properties, then the RK sub-step over
 * dt / 2 (src/StrangSplitting.cu:568-638).  The dt is fixed, or the CFL step of
 * local_time_step_without_reaction; returns the dt used. */
double oc_wu_step_blocks(oc_zone *const *zones, int n_block, const oc_run *r, const oc_tables *tb, double dt_fixed,
                         int fixed);

/* 0-D constant-volume reactor, explicit SSP-RK3 on rho*Y at fixed rho and e
 * (src/ConstVolumeReactor.cpp:708-745 RK3 branch, update_t 1456-1482).
 * y_inout holds mass fractions; t_inout the temperature. Returns steps taken. */
long oc_cv_reactor(const oc_tables *tb, double rho, double *t_inout, double *y_inout, double dt, double t_end,
                   double *trace_t, double *trace_T, long trace_every, long trace_cap);
/* The same reactor with the point-implicit diagonal update (scheme 0,
 * src/ConstVolumeReactor.cpp:400-440) or SSP-RK3 (scheme 3); conserve_mass
 * renormalises the mass fractions after the point-implicit update. */
long oc_cv_reactor_ex(const oc_tables *tb, double rho, double *t_inout, double *y_inout, double dt, double t_end,
                      int scheme, int conserve_mass);
/* ROK4E with the full chemistry Jacobian (src/ConstVolumeReactor.cpp:442-707),
 * first step h0, adaptive with RTol/ATol or fixed steps; returns accepted steps. */
long oc_cv_reactor_rok4e(const oc_tables *tb, double rho, double *t_inout, double *y_inout, double h0, double t_end,
                         int adaptive, double rtol, double atol, long *rejected_out);

#ifdef __cplusplus
}
#endif
#endif
