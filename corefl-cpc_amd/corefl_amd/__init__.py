"""MI355X (gfx950) implementation of COREFL's per-stage RHS and SSP-RK3 update.

The compute path is libcorefl_amd.so (hand-written HIP kernels behind a C ABI,
include/corefl_amd.h).  This package binds it for tests and benchmarks.
"""
from . import _abi
from .cases import Case, channel_case, metrics, multiblock_case, reacting_box_case, sod_case, face_list, H2_MECH
from .solver import CoreflError, Solver, load_tables, run_desc

__all__ = ["_abi", "Case", "channel_case", "metrics", "multiblock_case", "reacting_box_case", "sod_case", "face_list", "H2_MECH", "CoreflError",
           "Solver", "load_tables", "run_desc"]
