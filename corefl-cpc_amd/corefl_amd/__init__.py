"""MI355X (gfx950) implementation of COREFL's per-stage RHS and SSP-RK3 update.

The compute path is libcorefl_amd.so (hand-written HIP kernels behind a C ABI,
include/corefl_amd.h).  This package binds it for tests and benchmarks.
"""
from . import _abi
from .cases import (AIR5_MECH, H2_MECH, Case, air5_box_case, channel_case, face_list, metrics, multiblock_case,
                    reacting_box_case, sod_case, square_ogrid_case, tube_case)
from .serve import ReactorServer, ServeError, TubeServer, h2_air_mass_fractions
from .solver import CoreflError, Solver, load_air5_tables, load_tables, run_desc

__all__ = ["_abi", "Case", "air5_box_case", "channel_case", "metrics", "multiblock_case", "reacting_box_case",
           "sod_case", "square_ogrid_case", "face_list", "AIR5_MECH", "H2_MECH", "CoreflError", "Solver", "load_air5_tables", "load_tables",
           "run_desc", "ReactorServer", "ServeError", "TubeServer", "tube_case", "h2_air_mass_fractions"]
