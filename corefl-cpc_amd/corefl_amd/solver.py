"""Thin Python handle over the C ABI (tests and bench drive the library through it)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _abi
from .cases import AIR5_MECH, H2_MECH, Case


class CoreflError(RuntimeError):
    """A failed C-ABI call; status is the corefl_status it returned
    (-5, COREFL_ERR_STATE: a non-finite or non-positive state)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def load_tables(mech=None, therm=None, tran=None) -> _abi.ChemTables:
    """Chemkin tables through corefl_chem_load (defaults: Li-2004 H2/O2/N2)."""
    lib = _abi.load()
    mech = mech or H2_MECH / "mechanism.inp"
    therm = therm or H2_MECH / "therm.dat"
    tran = tran or H2_MECH / "tran.dat"
    t = _abi.ChemTables()
    st = lib.corefl_chem_load(str(mech).encode(), str(therm).encode(), str(tran).encode(), C.byref(t))
    if st != 0:
        raise CoreflError(f"corefl_chem_load failed ({st})")
    return t


def load_air5_tables(two_temperature: bool = True) -> _abi.ChemTables:
    """AIR-5 (Park's 17 reactions) with NASA-9 thermodynamics, and the
    two-temperature metadata unless two_temperature is False."""
    lib = _abi.load()
    t = _abi.ChemTables()
    files = [str(AIR5_MECH / f).encode() for f in ("mechanism.inp", "therm.dat", "tran.dat")]
    tt = str(AIR5_MECH / "two_temperature.dat").encode() if two_temperature else None
    st = lib.corefl_chem_load_ex(*files, 9, tt, C.byref(t))
    if st != 0:
        raise CoreflError(f"corefl_chem_load_ex failed ({st})")
    return t


# defaults for keys a run dict may leave out: the reference's (src/Parameter.cpp:393-396), and
# chem_kernel 0 (the chemistry kernel compiled for the mechanism); one temperature
RUN_DEFAULTS = dict(reconstruction=2, limiter=0, entropy_fix_factor=0.125, limit_flow=0, p_ref=1.0,
                    chem_kernel=0, two_temperature=0, vt_relaxation=0, positive_preserving=0, shock_sensor=0,
                    sensor_eps=-1.0, sensor_threshold=0.0, step_graph=0, autotune=1, temporal_scheme=3,
                    chem_advance_method=0)


def run_desc(d: dict) -> _abi.RunDesc:
    r = _abi.RunDesc()
    for k, v in {**RUN_DEFAULTS, **d}.items():
        setattr(r, k, v)
    return r


class Solver:
    """One context (one GPU) with blocks added from Case objects."""

    def __init__(self, run: dict | _abi.RunDesc, tables: _abi.ChemTables | None = None):
        self.lib = _abi.load()
        self.run = run if isinstance(run, _abi.RunDesc) else run_desc(run)
        self.ctx = C.c_void_p()
        st = self.lib.corefl_ctx_create(C.byref(self.run), C.byref(tables) if tables is not None else None,
                                        C.byref(self.ctx))
        if st != 0:
            msg = self.lib.corefl_last_error(self.ctx).decode() if self.ctx else "?"
            self.close()
            raise CoreflError(f"corefl_ctx_create failed ({st}): {msg}")
        self.cases: list[Case] = []

    def _check(self, st, what):
        if st != 0:
            raise CoreflError(f"{what} failed ({st}): {self.lib.corefl_last_error(self.ctx).decode()}", st)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.corefl_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bc_set(self, cond: dict):
        """Register the condition of one boundary label (corefl_bc_set)."""
        c = _abi.BcCondition()
        for k, v in cond.items():
            if k == "sv":
                c.sv[:len(v)] = [float(x) for x in v]
            else:
                setattr(c, k, v)
        self._check(self.lib.corefl_bc_set(self.ctx, C.byref(c)), "corefl_bc_set")

    def add_block(self, case: Case) -> int:
        for cond in case.conditions:
            self.bc_set(cond)
        faces = (_abi.Boundary * max(1, len(case.boundaries)))()
        for n, (face, direction, typ, start, end) in enumerate(case.boundaries):
            faces[n].face, faces[n].direction, faces[n].type = face, direction, typ
            for a in range(3):
                faces[n].range_start[a] = start[a]
                faces[n].range_end[a] = end[a]
        inner = (_abi.InnerFace * max(1, len(case.inner)))()
        for n, (src, tar, tblk) in enumerate(case.inner):
            inner[n].src[:] = list(src)
            inner[n].tar[:] = list(tar)
            inner[n].target_block = tblk
        metric = np.ascontiguousarray(case.metric, dtype=np.float64)
        jac = np.ascontiguousarray(case.jac, dtype=np.float64)
        d = _abi.BlockDesc(case.mx, case.my, case.mz, case.ngg, _dptr(metric), _dptr(jac), len(case.boundaries),
                           faces, len(case.inner), inner)
        bid = C.c_int()
        self._check(self.lib.corefl_block_add(self.ctx, C.byref(d), C.byref(bid)), "corefl_block_add")
        self.cases.append(case)
        self.upload(bid.value, case.bv, case.sv)
        if case.tve is not None:
            self.upload_field(bid.value, _abi.F_TVE, case.tve)
        return bid.value

    def upload(self, blk: int, bv: np.ndarray, sv: np.ndarray):
        bv = np.ascontiguousarray(bv, dtype=np.float64)
        sv = np.ascontiguousarray(sv, dtype=np.float64)
        self._check(self.lib.corefl_state_upload(self.ctx, blk, _dptr(bv), _dptr(sv) if sv.size else None),
                    "corefl_state_upload")

    def download(self, blk: int, field: int) -> np.ndarray:
        n = self.lib.corefl_field_size(self.ctx, blk, field)
        out = np.empty(n, dtype=np.float64)
        self._check(self.lib.corefl_field_download(self.ctx, blk, field, _dptr(out)), "corefl_field_download")
        c = self.cases[blk]
        if field in (_abi.F_DQ, _abi.F_DT_LOCAL):
            return out.reshape(-1, c.mz, c.my, c.mx)
        return out.reshape((-1,) + c.padded_shape)

    def upload_field(self, blk: int, field: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        n = self.lib.corefl_field_size(self.ctx, blk, field)
        if a.size != n:
            raise ValueError(f"field size {a.size} != {n}")
        self._check(self.lib.corefl_field_upload(self.ctx, blk, field, _dptr(a)), "corefl_field_upload")

    def initialize(self):
        self._check(self.lib.corefl_initialize(self.ctx), "corefl_initialize")

    def op(self, op: int, blk: int = -1, rk: int = 0, dt: float = 0.0):
        self._check(self.lib.corefl_op(self.ctx, op, blk, rk, dt), f"corefl_op({_abi.OP_NAMES[op]})")

    def stage(self, rk: int, dt: float):
        self._check(self.lib.corefl_stage(self.ctx, rk, dt), "corefl_stage")

    def step(self, dt: float = 0.0) -> float:
        v = C.c_double(dt)
        self._check(self.lib.corefl_step(self.ctx, C.byref(v)), "corefl_step")
        return v.value

    def sync(self):
        self._check(self.lib.corefl_sync(self.ctx), "corefl_sync")

    def tiles(self, blk: int = 0) -> list:
        """Tile shape (0 / 1) of each direction's characteristic WENO sweep on a block."""
        v = (C.c_int * 3)()
        self._check(self.lib.corefl_block_tiles(self.ctx, blk, v), "corefl_block_tiles")
        return list(v)

    def set_tiles(self, blk: int, shapes) -> None:
        v = (C.c_int * 3)(*shapes)
        self._check(self.lib.corefl_block_set_tiles(self.ctx, blk, v), "corefl_block_set_tiles")

    def cfl_dt(self) -> float:
        """Global CFL time step (local_time_step + min) on the current state."""
        self.op(_abi.OP_LOCAL_DT)
        dt = min(float(self.download(b, _abi.F_DT_LOCAL).min()) for b in range(len(self.cases)))
        return dt

    def write_plt(self, path, coords=None, ngg_out: int = 0, solution_time: float = 0.0, step_path=None,
                  step: int = 0):
        """Tecplot binary solution file of every block (corefl_plt_write, the reference's
        output/flowfield.plt).  coords: per block the node coordinates (3, nzp, nyp, nxp) over the
        padded extent; by default the cases' x where they carry one, else zeros."""
        nb = len(self.cases)
        if coords is None:
            coords = [c.x for c in self.cases]
        keep = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in coords]
        ptrs = (C.POINTER(C.c_double) * max(1, nb))()
        for b in range(nb):
            ptrs[b] = _dptr(keep[b]) if keep[b] is not None else C.POINTER(C.c_double)()
        self._check(self.lib.corefl_plt_write(self.ctx, str(path).encode(), ptrs, ngg_out, solution_time,
                                              str(step_path).encode() if step_path else None, step),
                    "corefl_plt_write")

    def read_plt(self, path) -> float:
        """Restart: the blocks' bv, sv (and Tve) from a solution file (corefl_plt_read); returns its
        solution time.  Run initialize() afterwards."""
        t = C.c_double(0.0)
        self._check(self.lib.corefl_plt_read(self.ctx, str(path).encode(), C.byref(t)), "corefl_plt_read")
        return t.value

    def write_output(self, case_dir, step: int, solution_time: float = 0.0, ngg_out: int = 0):
        """case_dir/output/flowfield.plt and case_dir/output/message/step.txt, as the reference's
        FieldIO<>::print_field leaves them (src/FieldIO.h:709-800)."""
        out = Path(case_dir) / "output"
        (out / "message").mkdir(parents=True, exist_ok=True)
        self.write_plt(out / "flowfield.plt", ngg_out=ngg_out, solution_time=solution_time,
                       step_path=out / "message" / "step.txt", step=step)

    def restart(self, case_dir):
        """The reference's restart (read_flowfield, src/Initialize.cu:167-420): the state from
        case_dir/output/flowfield.plt, then initialize(); returns (solution_time, step) with the step
        from output/message/step.txt (0 when it is missing)."""
        out = Path(case_dir) / "output"
        t = self.read_plt(out / "flowfield.plt")
        self.initialize()
        step_file = out / "message" / "step.txt"
        step = int(step_file.read_text().split()[0]) if step_file.exists() else 0
        return t, step

    def profile_begin(self):
        self._check(self.lib.corefl_profile_begin(self.ctx), "corefl_profile_begin")

    def profile_end(self):
        ms = (C.c_double * _abi.OP_COUNT)()
        calls = (C.c_long * _abi.OP_COUNT)()
        self._check(self.lib.corefl_profile_end(self.ctx, ms, calls), "corefl_profile_end")
        return {_abi.OP_NAMES[i]: (ms[i], calls[i]) for i in range(_abi.OP_COUNT)}
