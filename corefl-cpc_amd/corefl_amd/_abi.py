"""ctypes mirror of include/corefl_amd.h and loader for libcorefl_amd.so.

The product is the HIP library; this module only binds it.  Loading fails
loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

MAX_SPEC = 20
SPEC_NAME_LEN = 24  # COREFL_SPEC_NAME_LEN
MAX_REAC = 64
MAX_RANGE = 4
MAX_ELEC = 16

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libcorefl_amd.so"

S, R, NR, NE = MAX_SPEC, MAX_REAC, MAX_RANGE, MAX_ELEC
D, I = C.c_double, C.c_int


class ChemTables(C.Structure):
    _fields_ = [
        ("n_spec", I), ("n_reac", I), ("nasa7or9", I),
        ("mw", D * S), ("imw", D * S), ("gas_const", D * S),
        ("t_low", D * S), ("t_mid", D * S), ("t_high", D * S),
        ("low_coeff", (D * 7) * S), ("high_coeff", (D * 7) * S),
        ("geometry", I * S),
        ("vis_coeff", D * S), ("LJ_potent_inv", D * S), ("ZRotF298", D * S),
        ("WjDivWi_to_One4th", (D * S) * S), ("sqrt_WiDivWjPl1Mul8", (D * S) * S),
        ("kb_over_eps_jk", (D * S) * S), ("binary_diffusivity_coeff", (D * S) * S),
        ("reac_type", I * R), ("rev_type", I * R), ("order", I * R),
        ("stoi_f", (I * S) * R), ("stoi_b", (I * S) * R),
        ("A", D * R), ("b", D * R), ("Ea", D * R), ("A2", D * R), ("b2", D * R), ("Ea2", D * R),
        ("third_body", (D * S) * R),
        ("troe_alpha", D * R), ("troe_t1", D * R), ("troe_t2", D * R), ("troe_t3", D * R),
        ("n_range", I * S), ("t_cut", (D * (NR + 1)) * S), ("poly9", ((D * 9) * NR) * S),
        ("two_temperature", I), ("theta_v", D * S), ("n_elec", I * S),
        ("elec_theta", (D * NE) * S), ("elec_g", (D * NE) * S),
        ("lt_a", (D * S) * S), ("lt_b", (D * S) * S), ("park_sigma", D * S),
        ("trx_model", I), ("tcf_a", D * R), ("tcf_b", D * R), ("tcb_a", D * R), ("tcb_b", D * R),
        ("species_name", (C.c_char * SPEC_NAME_LEN) * S),
    ]


class RunDesc(C.Structure):
    _fields_ = [
        ("mixture", I), ("n_spec", I), ("n_var", I), ("n_scalar", I), ("n_scalar_transported", I),
        ("inviscid_scheme", I), ("viscous_order", I), ("reaction", I), ("gradp_in_diffusion", I),
        ("fixed_time_step", I), ("dt", D), ("cfl", D), ("prandtl", D), ("rho_ref", D), ("v_ref", D),
        ("reconstruction", I), ("limiter", I), ("entropy_fix_factor", D),
        ("limit_flow", I), ("p_ref", D), ("chem_kernel", I), ("two_temperature", I), ("vt_relaxation", I),
        ("positive_preserving", I), ("shock_sensor", I), ("sensor_eps", D), ("sensor_threshold", D),
        ("step_graph", I), ("autotune", I), ("temporal_scheme", I), ("chem_advance_method", I),
    ]


class Boundary(C.Structure):
    _fields_ = [("face", I), ("direction", I), ("type", I), ("range_start", I * 3), ("range_end", I * 3)]


class BcCondition(C.Structure):
    _fields_ = [("label", I), ("kind", I), ("thermal_type", I), ("inflow_sub", I), ("wall_temperature", D),
                ("density", D), ("u", D), ("v", D), ("w", D), ("pressure", D), ("temperature", D), ("mw", D),
                ("sv", D * MAX_SPEC), ("temperature_ve", D)]


class ReactorRequest(C.Structure):
    _fields_ = [("temperature", D), ("pressure", D), ("y", D * MAX_SPEC), ("t_end", D), ("dt", D), ("scheme", I),
                ("adaptive", I), ("rtol", D), ("atol", D), ("conserve_mass", I), ("max_steps", C.c_long)]


class ReactorResult(C.Structure):
    _fields_ = [("temperature", D), ("pressure", D), ("time", D), ("y", D * MAX_SPEC), ("steps", C.c_long),
                ("rejected", C.c_long), ("status", I), ("reason", I)]


class TubeRequest(C.Structure):
    _fields_ = [("n_cells", I), ("n_left", I), ("length", D), ("t_end", D), ("bc_left", I), ("bc_right", I),
                ("temperature", D * 2), ("pressure", D * 2), ("velocity", D * 2), ("y", (D * MAX_SPEC) * 2),
                ("max_steps", C.c_long)]


class TubeResult(C.Structure):
    _fields_ = [("time", D), ("steps", C.c_long), ("n_cells", I), ("status", I), ("reason", I)]


class InnerFace(C.Structure):
    _fields_ = [("src", I * 6), ("tar", I * 6), ("target_block", I)]


class BlockDesc(C.Structure):
    _fields_ = [
        ("mx", I), ("my", I), ("mz", I), ("ngg", I),
        ("metric", C.POINTER(D)), ("jac", C.POINTER(D)),
        ("n_boundary", I), ("boundary", C.POINTER(Boundary)),
        ("n_inner", I), ("inner", C.POINTER(InnerFace)),
    ]


# field ids (corefl_field)
F_CV, F_BV, F_SV, F_QN, F_DQ, F_FFLUX, F_GFLUX, F_HFLUX, F_METRIC, F_JAC, F_MUL, F_KAPPA, F_GAMMA, \
    F_ACOUSTIC, F_MACH, F_RHOD, F_RTS, F_DT_LOCAL, F_TVE, F_KAPPA_VE, F_SHOCK_SENSOR = range(21)
# operator ids (corefl_op_id)
OP_ZERO_DQ, OP_VISCOUS, OP_CHEMISTRY, OP_INVISCID_X, OP_INVISCID_Y, OP_INVISCID_Z, OP_UPDATE_RK, OP_BC, \
    OP_K_GRADIENT, OP_PROPERTIES, OP_CV_FROM_BV, OP_LOCAL_DT, OP_COPY_QN, OP_HALO, OP_LIMIT_FLOW, \
    OP_SHOCK_SENSOR, OP_WU_CN, OP_WU_REACTION, OP_UPDATE_RK_WU = range(19)
OP_COUNT = 19
OP_NAMES = ["zero_dq", "viscous", "chemistry", "inviscid_x", "inviscid_y", "inviscid_z", "update_rk", "bc",
            "k_gradient", "properties", "cv_from_bv", "local_dt", "copy_qn", "halo", "limit_flow", "shock_sensor",
            "wu_cn", "wu_reaction", "update_rk_wu"]

# symbols declared in include/corefl_amd.h
EXPORTS = [
    "corefl_chem_load", "corefl_chem_load_ex", "corefl_setup_load", "corefl_ctx_create", "corefl_ctx_destroy", "corefl_last_error",
    "corefl_bc_set", "corefl_block_add", "corefl_state_upload", "corefl_field_download", "corefl_field_upload",
    "corefl_field_size", "corefl_initialize", "corefl_op", "corefl_stage", "corefl_step", "corefl_sync",
    "corefl_profile_begin", "corefl_profile_end", "corefl_chem_kernel_check", "corefl_block_tiles",
    "corefl_block_set_tiles",
    "corefl_serve_open", "corefl_serve_submit", "corefl_serve_progress", "corefl_serve_poll", "corefl_serve_drain",
    "corefl_serve_capacity", "corefl_serve_error", "corefl_serve_close", "corefl_serve_run_noalloc",
    "corefl_tubes_open", "corefl_tubes_submit", "corefl_tubes_progress", "corefl_tubes_poll", "corefl_tubes_ready",
    "corefl_tubes_capacity", "corefl_tubes_error", "corefl_tubes_close",
    "corefl_mesh_load", "corefl_mesh_n_block", "corefl_mesh_block", "corefl_mesh_coords", "corefl_mesh_initial_state",
    "corefl_mesh_conditions", "corefl_mesh_error", "corefl_mesh_free",
    "corefl_plt_write", "corefl_plt_read",
]

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libcorefl_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python corefl-cpc_amd/build.py` "
                           "(the HIP library is the only implementation)")
    lib = C.CDLL(str(p))
    vp = C.c_void_p
    lib.corefl_chem_load.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(ChemTables)]
    lib.corefl_chem_load_ex.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_char_p,
                                        C.POINTER(ChemTables)]
    lib.corefl_chem_kernel_check.argtypes = [C.POINTER(ChemTables), C.c_char_p, C.c_size_t]
    lib.corefl_setup_load.argtypes = [C.c_char_p, C.POINTER(RunDesc), C.POINTER(ChemTables)]
    lib.corefl_ctx_create.argtypes = [C.POINTER(RunDesc), C.POINTER(ChemTables), C.POINTER(vp)]
    lib.corefl_block_tiles.argtypes = [vp, I, C.POINTER(I)]
    lib.corefl_block_set_tiles.argtypes = [vp, I, C.POINTER(I)]
    lib.corefl_ctx_destroy.argtypes = [vp]
    lib.corefl_ctx_destroy.restype = None
    lib.corefl_last_error.argtypes = [vp]
    lib.corefl_last_error.restype = C.c_char_p
    lib.corefl_bc_set.argtypes = [vp, C.POINTER(BcCondition)]
    lib.corefl_block_add.argtypes = [vp, C.POINTER(BlockDesc), C.POINTER(I)]
    lib.corefl_state_upload.argtypes = [vp, I, C.POINTER(D), C.POINTER(D)]
    lib.corefl_field_download.argtypes = [vp, I, I, C.POINTER(D)]
    lib.corefl_field_upload.argtypes = [vp, I, I, C.POINTER(D)]
    lib.corefl_field_size.argtypes = [vp, I, I]
    lib.corefl_field_size.restype = C.c_long
    lib.corefl_initialize.argtypes = [vp]
    lib.corefl_op.argtypes = [vp, I, I, I, D]
    lib.corefl_stage.argtypes = [vp, I, D]
    lib.corefl_step.argtypes = [vp, C.POINTER(D)]
    lib.corefl_sync.argtypes = [vp]
    lib.corefl_profile_begin.argtypes = [vp]
    lib.corefl_profile_end.argtypes = [vp, C.POINTER(D), C.POINTER(C.c_long)]
    lib.corefl_serve_open.argtypes = [C.POINTER(ChemTables), C.c_size_t, I, C.POINTER(vp)]
    lib.corefl_serve_submit.argtypes = [vp, C.POINTER(ReactorRequest)]
    lib.corefl_serve_submit.restype = C.c_long
    lib.corefl_serve_progress.argtypes = [vp, C.POINTER(I)]
    lib.corefl_serve_poll.argtypes = [vp, C.c_long, C.POINTER(ReactorResult)]
    lib.corefl_serve_drain.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(ReactorResult), I]
    lib.corefl_serve_capacity.argtypes = [vp]
    lib.corefl_serve_error.argtypes = [vp]
    lib.corefl_serve_error.restype = C.c_char_p
    lib.corefl_serve_close.argtypes = [vp]
    lib.corefl_serve_close.restype = None
    lib.corefl_serve_run_noalloc.argtypes = [C.POINTER(ChemTables), C.POINTER(ReactorRequest), I, I,
                                             C.POINTER(ReactorResult)]
    lib.corefl_tubes_open.argtypes = [C.POINTER(RunDesc), C.POINTER(ChemTables), C.c_size_t, I, C.POINTER(vp)]
    lib.corefl_tubes_submit.argtypes = [vp, C.POINTER(TubeRequest)]
    lib.corefl_tubes_submit.restype = C.c_long
    lib.corefl_tubes_progress.argtypes = [vp, C.POINTER(I)]
    lib.corefl_tubes_poll.argtypes = [vp, C.c_long, C.POINTER(TubeResult), C.POINTER(D), C.c_long]
    lib.corefl_tubes_ready.argtypes = [vp, C.POINTER(C.c_long), I]
    lib.corefl_tubes_capacity.argtypes = [vp]
    lib.corefl_tubes_capacity.restype = C.c_long
    lib.corefl_tubes_error.argtypes = [vp]
    lib.corefl_tubes_error.restype = C.c_char_p
    lib.corefl_tubes_close.argtypes = [vp]
    lib.corefl_tubes_close.restype = None
    lib.corefl_mesh_load.argtypes = [C.c_char_p, I, C.POINTER(vp)]
    lib.corefl_mesh_n_block.argtypes = [vp]
    lib.corefl_mesh_block.argtypes = [vp, I, C.POINTER(BlockDesc)]
    lib.corefl_mesh_coords.argtypes = [vp, I, C.POINTER(D)]
    lib.corefl_mesh_initial_state.argtypes = [vp, I, C.POINTER(ChemTables), C.POINTER(D), C.POINTER(D), C.POINTER(D)]
    lib.corefl_mesh_conditions.argtypes = [vp, C.POINTER(ChemTables), C.POINTER(BcCondition), I]
    lib.corefl_mesh_error.argtypes = [vp]
    lib.corefl_mesh_error.restype = C.c_char_p
    lib.corefl_mesh_free.argtypes = [vp]
    lib.corefl_mesh_free.restype = None
    lib.corefl_plt_write.argtypes = [vp, C.c_char_p, C.POINTER(C.POINTER(D)), I, D, C.c_char_p, I]
    lib.corefl_plt_read.argtypes = [vp, C.c_char_p, C.POINTER(D)]
    if path is None:
        _lib = lib
    return lib
