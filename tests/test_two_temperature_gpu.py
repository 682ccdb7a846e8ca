"""Parity of the two-temperature path (5-species air with E_ve, NASA-9,
Park's control temperature, V-T relaxation) on the GPU with the oracle, on a
curvilinear periodic box with Tve between 0.5 T and 1.5 T: initialisation
(E_ve from Tve), every right-hand-side operator for the characteristic WENO
and the primitive-variable solvers, the 2nd- and 8th-order viscous terms with
the kappa_ve and E_ve diffusion terms, a full stage and ten steps."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, air5_box_case, load_air5_tables
from oracle import Oracle, air5_tables

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-11
LONG_TOL = 1e-9
FIELDS = [("cv", _abi.F_CV), ("bv", _abi.F_BV), ("sv", _abi.F_SV)]


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
    return arr[..., g:g + case.mz, g:g + case.my, g:g + case.mx]


def check(name, gpu, ref, tol, info=""):
    err = normwise(gpu, ref, vector=(1, 2, 3) if name in ("cv", "bv") else None)
    assert np.all(err <= tol), f"{name}{info}: normwise error per variable {err} > {tol}"


@pytest.fixture(scope="module")
def tables():
    return load_air5_tables(two_temperature=True)


@pytest.fixture(scope="module")
def oc_tables():
    return air5_tables(two_temperature=True)


def pair(case, tables, oc_tables):
    r = case.run
    print(f"air5 2T scheme {r['inviscid_scheme']} viscous {r['viscous_order']}", flush=True)
    s = Solver(r, tables)
    s.add_block(case)
    o = Oracle(case, oc_tables)
    s.initialize()
    o.initialize()
    return s, o


def check_state(case, s, o, tol, info):
    for name, fid in FIELDS:
        check(name, interior(case, s.download(0, fid)), interior(case, o.a[name]), tol, info)
    check("tve", interior(case, s.download(0, _abi.F_TVE)), interior(case, o.a["tve"][None]), tol, info)


def test_two_temperature_initialize_parity(tables, oc_tables):
    case = air5_box_case(20, 16, 12, seed=21, tables=tables, two_temperature=True)
    s, o = pair(case, tables, oc_tables)
    check_state(case, s, o, STAGE_TOL, " after initialize")
    for name, fid in (("kappa_ve", _abi.F_KAPPA_VE), ("gamma", _abi.F_GAMMA), ("acoustic", _abi.F_ACOUSTIC)):
        check(name, interior(case, s.download(0, fid)), interior(case, o.a[name][None]), STAGE_TOL)
    s.close()


@pytest.mark.parametrize("scheme", [52, 3, 4, 2])
@pytest.mark.parametrize("op", ["inviscid_x", "inviscid_y", "inviscid_z", "viscous", "chemistry"])
def test_two_temperature_operator_parity(tables, oc_tables, scheme, op):
    if op in ("viscous", "chemistry") and scheme != 52:
        pytest.skip("scheme-independent operator")
    case = air5_box_case(20, 16, 12, seed=22, tables=tables, two_temperature=True)
    case.run.update(inviscid_scheme=scheme, reconstruction=2)
    s, o = pair(case, tables, oc_tables)
    s.op(_abi.OP_ZERO_DQ)
    o.op("zero_dq")
    s.op(getattr(_abi, "OP_" + op.upper()))
    o.op(op)
    check("dq", s.download(0, _abi.F_DQ), o.a["dq"], STAGE_TOL, f" after {op}")
    s.close()


@pytest.mark.parametrize("scheme", [52, 72, 3])
def test_two_temperature_stage_and_steps_parity(tables, oc_tables, scheme):
    case = air5_box_case(20, 16, 12, seed=23, tables=tables, two_temperature=True, ngg=4 if scheme == 72 else 3)
    case.run.update(inviscid_scheme=scheme)
    s, o = pair(case, tables, oc_tables)
    dt = 0.5 * o.cfl_dt()
    s.op(_abi.OP_COPY_QN)
    o.op("copy_qn")
    for rk in range(3):
        s.stage(rk, dt)
        o.stage(rk, dt)
        check_state(case, s, o, STAGE_TOL, f" after stage {rk}")
    for _ in range(10):
        s.step(dt)
        o.step(dt, fixed=True)
    check_state(case, s, o, LONG_TOL, " after 11 steps")
    y = interior(case, s.download(0, _abi.F_SV))[:5]
    assert y.min() >= 0.0
    assert np.abs(y.sum(axis=0) - 1.0).max() <= 1e-12
    s.close()


def test_two_temperature_viscous8_parity(tables, oc_tables):
    """8th-order collocated viscous terms with kappa_ve grad Tve, the
    non-equilibrium diffusion enthalpy and the E_ve flux
    (src/ViscousScheme.cu:470-834): the operator, a stage and ten steps."""
    case = air5_box_case(20, 16, 12, seed=25, tables=tables, two_temperature=True, ngg=4)
    case.run.update(viscous_order=8, inviscid_scheme=72)
    s, o = pair(case, tables, oc_tables)
    s.op(_abi.OP_ZERO_DQ)
    o.op("zero_dq")
    s.op(_abi.OP_VISCOUS)
    o.op("viscous")
    check("dq", s.download(0, _abi.F_DQ), o.a["dq"], STAGE_TOL, " after viscous (8th order)")
    dt = 0.5 * o.cfl_dt()
    s.op(_abi.OP_COPY_QN)
    o.op("copy_qn")
    for rk in range(3):
        s.stage(rk, dt)
        o.stage(rk, dt)
        check_state(case, s, o, STAGE_TOL, f" after stage {rk}")
    for _ in range(10):
        s.step(dt)
        o.step(dt, fixed=True)
    check_state(case, s, o, LONG_TOL, " after 11 steps")
    s.close()
