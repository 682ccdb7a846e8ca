"""Reference case directories through the C ABI (corefl_setup_load and the
corefl_mesh_* readers of include/corefl_amd.h): setup.txt, the readGrid
output grid/grid%4d.dat|grd and boundary_condition/{boundary,inner}%4d.txt
under case_dir/input/.  load_case_dir turns them into Case objects (metric,
Jacobian, boundary faces, interfaces, the conditions of the setup's
boundary structs and the grouped initial state), which the Solver and the
oracle both take.  write_grid / write_boundary write the same formats."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _abi
from .cases import Case


class CaseDirError(RuntimeError):
    pass


def _rank_name(prefix: str, rank: int, ext: str) -> str:
    return f"{prefix}{rank:4d}{ext}"


def write_grid(input_dir, blocks, rank: int = 0, binary: bool = True) -> Path:
    """blocks: list of (X, Y, Z) node arrays, each shaped (mz, my, mx)."""
    d = Path(input_dir) / "grid"
    d.mkdir(parents=True, exist_ok=True)
    if binary:
        p = d / _rank_name("grid", rank, ".dat")
        with open(p, "wb") as f:
            np.array([len(blocks)], dtype=np.int32).tofile(f)
            for X, _, _ in blocks:
                mz, my, mx = X.shape
                np.array([mx, my, mz], dtype=np.int32).tofile(f)
            for X, Y, Z in blocks:
                for A in (X, Y, Z):
                    np.ascontiguousarray(A, dtype=np.float64).tofile(f)  # i fastest
    else:
        p = d / _rank_name("grid", rank, ".grd")
        with open(p, "w") as f:
            f.write(f"{len(blocks)}\n")
            for X, _, _ in blocks:
                mz, my, mx = X.shape
                f.write(f"{mx} {my} {mz}\n")
            for X, Y, Z in blocks:
                for A in (X, Y, Z):
                    f.write("\n".join(repr(float(v)) for v in A.ravel()) + "\n")
    return p


def write_boundary(input_dir, faces_per_block, rank: int = 0) -> Path:
    """faces_per_block: per block a list of (i1, i2, j1, j2, k1, k2, label), 0-based."""
    d = Path(input_dir) / "boundary_condition"
    d.mkdir(parents=True, exist_ok=True)
    p = d / _rank_name("boundary", rank, ".txt")
    with open(p, "w") as f:
        for faces in faces_per_block:
            f.write(f"{len(faces):7d}\n")
            for fc in faces:
                f.write("".join(f"{v:7d}" for v in fc) + "\n")
    return p


def _run_dict(rd: _abi.RunDesc) -> dict:
    return {name: getattr(rd, name) for name, _ in rd._fields_}


class CaseDir:
    """A loaded case directory: run switches, tables and one Case per block."""

    def __init__(self, case_dir, rank: int = 0):
        self.lib = _abi.load()
        self.dir = str(case_dir)
        rd, tables = _abi.RunDesc(), _abi.ChemTables()
        st = self.lib.corefl_setup_load(self.dir.encode(), C.byref(rd), C.byref(tables))
        if st != 0:
            raise CaseDirError(f"corefl_setup_load failed ({st}) for {self.dir}")
        self.run = _run_dict(rd)
        self.tables = tables if rd.mixture else None
        self.h = C.c_void_p()
        st = self.lib.corefl_mesh_load(self.dir.encode(), rank, C.byref(self.h))
        if st != 0:
            msg = self.lib.corefl_mesh_error(self.h).decode() if self.h else ""
            self.close()
            raise CaseDirError(f"corefl_mesh_load failed ({st}): {msg}")
        tb = C.byref(tables) if self.tables is not None else None
        n = self.lib.corefl_mesh_conditions(self.h, tb, None, 0)
        if n < 0:
            raise CaseDirError(f"corefl_mesh_conditions failed: {self.lib.corefl_mesh_error(self.h).decode()}")
        conds = (_abi.BcCondition * max(1, n))()
        self.lib.corefl_mesh_conditions(self.h, tb, conds, n)
        ns = tables.n_spec if self.tables is not None else 0
        self.conditions = []
        for c in conds[:n]:
            d = {name: getattr(c, name) for name, _ in c._fields_ if name != "sv"}
            d["sv"] = [c.sv[l] for l in range(ns + (1 if rd.two_temperature else 0))]
            self.conditions.append(d)
        self.cases = [self._block(b, ns) for b in range(self.lib.corefl_mesh_n_block(self.h))]

    def _block(self, b: int, ns: int) -> Case:
        d = _abi.BlockDesc()
        if self.lib.corefl_mesh_block(self.h, b, C.byref(d)) != 0:
            raise CaseDirError("corefl_mesh_block failed")
        g = d.ngg
        shp = (d.mz + 2 * g, d.my + 2 * g, d.mx + 2 * g)
        sz = int(np.prod(shp))
        metric = np.ctypeslib.as_array(d.metric, shape=(9 * sz,)).copy().reshape((9,) + shp)
        jac = np.ctypeslib.as_array(d.jac, shape=(sz,)).copy().reshape(shp)
        xyz = np.zeros((3,) + shp)
        self.lib.corefl_mesh_coords(self.h, b, xyz.ctypes.data_as(C.POINTER(C.c_double)))
        faces = []
        for q in range(d.n_boundary):
            f = d.boundary[q]
            faces.append((f.face, f.direction, f.type, list(f.range_start), list(f.range_end)))
        inner = []
        for q in range(d.n_inner):
            f = d.inner[q]
            inner.append((list(f.src), list(f.tar), f.target_block))
        tt = self.run["two_temperature"]
        nsc = ns + (1 if tt else 0)
        bv = np.zeros((6,) + shp)
        sv = np.zeros((max(nsc, 0),) + shp)
        tve = np.zeros(shp) if tt else None
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None
        tb = C.byref(self.tables) if self.tables is not None else None
        st = self.lib.corefl_mesh_initial_state(self.h, b, tb, P(bv), P(sv) if nsc else None, P(tve))
        if st != 0:
            raise CaseDirError(f"corefl_mesh_initial_state failed: {self.lib.corefl_mesh_error(self.h).decode()}")
        run = dict(self.run)
        return Case(d.mx, d.my, d.mz, g, metric, jac, bv, sv, faces, run, xyz, inner,
                    conditions=list(self.conditions), tve=tve)

    def close(self):
        if getattr(self, "h", None):
            self.lib.corefl_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_case_dir(case_dir, rank: int = 0) -> CaseDir:
    return CaseDir(case_dir, rank)
