"""Parity of the 5-species air path with NASA-9 thermodynamics (Park's 17
reactions, WENO5 characteristic, 2nd-order viscous terms) on the GPU with the
oracle: initialisation, each right-hand-side operator, a full stage and ten
steps, on a curvilinear periodic box crossing the 1000 K and 6000 K range cuts."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, air5_box_case, load_air5_tables
from oracle import Oracle, air5_tables

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-11
LONG_TOL = 1e-9
FIELDS = [("cv", _abi.F_CV), ("bv", _abi.F_BV), ("sv", _abi.F_SV)]
PROPS = [("mul", _abi.F_MUL), ("kappa", _abi.F_KAPPA), ("gamma", _abi.F_GAMMA), ("acoustic", _abi.F_ACOUSTIC),
         ("rhoD", _abi.F_RHOD)]


def normwise(a, b, vector=None):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    den = np.abs(b).max(axis=1)
    if vector is not None:
        den[list(vector)] = den[list(vector)].max()
    den = np.maximum(den, 1e-300)
    return np.abs(a - b).max(axis=1) / den


def interior(case, arr):
    g = case.ngg
    return arr[:, g:g + case.mz, g:g + case.my, g:g + case.mx]


def check(name, gpu, ref, tol, info=""):
    err = normwise(gpu, ref, vector=(1, 2, 3) if name in ("cv", "bv") else None)
    assert np.all(err <= tol), f"{name}{info}: normwise error per variable {err} > {tol}"


@pytest.fixture(scope="module")
def tables():
    return load_air5_tables(two_temperature=False)


@pytest.fixture(scope="module")
def oc_tables():
    return air5_tables(two_temperature=False)


def pair(case, tables, oc_tables):
    print(f"air5 nasa9 {case.mx}x{case.my}x{case.mz} viscous {case.run['viscous_order']}", flush=True)
    s = Solver(case.run, tables)
    s.add_block(case)
    o = Oracle(case, oc_tables)
    s.initialize()
    o.initialize()
    return s, o


def test_air5_initialize_parity(tables, oc_tables):
    case = air5_box_case(20, 16, 12, seed=5, tables=tables)
    s, o = pair(case, tables, oc_tables)
    for name, fid in FIELDS + PROPS:
        ref = o.a[name] if o.a[name].ndim == 4 else o.a[name][None]
        check(name, interior(case, s.download(0, fid)), interior(case, ref), STAGE_TOL)
    s.close()


@pytest.mark.parametrize("op", ["viscous", "chemistry", "inviscid_x", "inviscid_y", "inviscid_z"])
def test_air5_operator_parity(tables, oc_tables, op):
    case = air5_box_case(20, 16, 12, seed=6, tables=tables)
    s, o = pair(case, tables, oc_tables)
    s.op(_abi.OP_ZERO_DQ)
    o.op("zero_dq")
    s.op(getattr(_abi, "OP_" + op.upper()))
    o.op(op)
    check("dq", s.download(0, _abi.F_DQ), o.a["dq"], STAGE_TOL, f" after {op}")
    s.close()


def test_air5_stage_and_steps_parity(tables, oc_tables):
    case = air5_box_case(20, 16, 12, seed=7, tables=tables)
    s, o = pair(case, tables, oc_tables)
    dt = 0.5 * o.cfl_dt()
    s.op(_abi.OP_COPY_QN)
    o.op("copy_qn")
    for rk in range(3):
        s.stage(rk, dt)
        o.stage(rk, dt)
        for name, fid in FIELDS:
            check(name, s.download(0, fid), o.a[name], STAGE_TOL, f" after stage {rk}")
    for _ in range(10):
        s.step(dt)
        o.step(dt, fixed=True)
    for name, fid in FIELDS:
        check(name, s.download(0, fid), o.a[name], LONG_TOL, " after 11 steps")
    y = interior(case, s.download(0, _abi.F_SV))
    assert y.min() >= 0.0
    assert np.abs(y.sum(axis=0) - 1.0).max() <= 1e-12
    s.close()
