"""Parity of the WENO options on the GPU with the oracle: component-wise WENO5
/ WENO7 (inviscid_scheme 51 / 71) for the Air perfect gas, the 9-species
H2/air mixture and two-temperature AIR-5; the species positivity limiter
(component-wise in x, y, z and characteristic in x, with a fixed step and
with the local CFL step); the three shock sensors and the hybrid linear/WENO
switch; and ten full steps of the hybrid, limited scheme."""
import numpy as np
import pytest

from corefl_amd import (Solver, _abi, air5_box_case, load_air5_tables, load_tables, multiblock_case,
                        reacting_box_case)
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


def check(name, gpu, ref, tol, info=""):
    err = normwise(gpu, ref, vector=(1, 2, 3) if name in ("cv", "bv", "dq") else None)
    assert np.all(err <= tol), f"{name}{info}: normwise error per variable {err} > {tol}"


def interior(case, arr):
    g = case.ngg
    return arr[..., g:g + case.mz, g:g + case.my, g:g + case.mx]


@pytest.fixture(scope="module")
def h2():
    return load_tables()


@pytest.fixture(scope="module")
def air5():
    return load_air5_tables(two_temperature=True), air5_tables(two_temperature=True)


def species_front_case(tables, seed=11):
    """H2/air box whose mixture fraction is thresholded to 0 / 1, so that the
    high-order species fluxes overshoot next to the fronts."""
    c = reacting_box_case(16, 12, 10, seed=seed, tables=tables)
    z = (c.sv[0] > 0.5).astype(float)
    c.sv[:] = 0.0
    c.sv[0] = z
    c.sv[2] = 0.233 * (1 - z)
    c.sv[8] = 0.767 * (1 - z)
    mw = np.array(tables.mw[:9])
    imw = (c.sv / mw[:, None, None, None]).sum(axis=0)
    c.bv[0] = c.bv[4] / (8314.462618 * imw * c.bv[5])
    c.run.update(viscous_order=0, reaction=0)
    return c


def make(physics, scheme, h2, air5):
    if physics == "air":
        c = multiblock_case(split=(1, 1, 1), widths=[[20], [16], [12]], seed=6, physics="air", rotate=False,
                            ngg=4)[0]
        tb = otb = None
    elif physics == "h2":
        c = reacting_box_case(20, 16, 12, seed=31, tables=h2)
        c.run.update(viscous_order=0, reaction=0)
        tb, otb = h2, None
    else:
        tb, otb = air5
        c = air5_box_case(20, 16, 12, seed=32, tables=tb, two_temperature=True, ngg=4)
    c.run.update(inviscid_scheme=scheme)
    return c, tb, otb


def pair(case, tables, oracle_tables, h2_oracle_tables=None):
    r = case.run
    print(f"scheme {r['inviscid_scheme']} mixture {r['mixture']} pp {r.get('positive_preserving', 0)} "
          f"sensor {r.get('shock_sensor', 0)}/{r.get('sensor_threshold', 0.0)} fixed {r.get('fixed_time_step')}",
          flush=True)
    s = Solver(r, tables)
    s.add_block(case)
    o = Oracle(case, oracle_tables if oracle_tables is not None else h2_oracle_tables)
    s.initialize()
    o.initialize()
    return s, o


@pytest.mark.parametrize("physics", ["air", "h2", "air5"])
@pytest.mark.parametrize("scheme", [51, 71])
def test_component_wise_operator_parity(h2, air5, h2_oracle_tables, physics, scheme):
    case, tb, otb = make(physics, scheme, h2, air5)
    s, o = pair(case, tb, otb, h2_oracle_tables if physics == "h2" else None)
    for d in "xyz":
        s.op(_abi.OP_ZERO_DQ)
        o.op("zero_dq")
        s.op(getattr(_abi, "OP_INVISCID_" + d.upper()))
        o.op("inviscid_" + d)
        check("dq", s.download(0, _abi.F_DQ), o.a["dq"], STAGE_TOL, f" after inviscid_{d}")
    s.close()


@pytest.mark.parametrize("scheme", [71, 51, 72])
@pytest.mark.parametrize("fixed", [1, 0])
def test_positivity_limiter_parity(h2, h2_oracle_tables, scheme, fixed):
    case = species_front_case(h2)
    case.run.update(inviscid_scheme=scheme, fixed_time_step=1)
    o0 = Oracle(case, h2_oracle_tables)
    o0.initialize()
    dt = 1.5 * o0.cfl_dt()
    case.run.update(positive_preserving=1, fixed_time_step=fixed, dt=dt if fixed else 0.0, cfl=0.75)
    s, o = pair(case, h2, h2_oracle_tables)
    if not fixed:
        s.op(_abi.OP_LOCAL_DT)
        o.cfl_dt()
        check("dt_local", s.download(0, _abi.F_DT_LOCAL)[None], o.a["dt_local"][None], STAGE_TOL)
    acted = 0
    for d in "xyz":
        s.op(_abi.OP_ZERO_DQ)
        o.op("zero_dq")
        s.op(getattr(_abi, "OP_INVISCID_" + d.upper()))
        o.op("inviscid_" + d)
        check("dq", s.download(0, _abi.F_DQ), o.a["dq"], STAGE_TOL, f" after inviscid_{d}")
        if d == "x":
            # the limiter changed fluxes here (same run without it differs)
            case.run.update(positive_preserving=0)
            off = Oracle(case, h2_oracle_tables)
            off.initialize()
            if not fixed:
                off.cfl_dt()
            off.op("zero_dq")
            off.op("inviscid_x")
            acted = int((off.a["ff"] != o.a["ff"]).sum())
            case.run.update(positive_preserving=1)
    assert acted > 0
    s.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_shock_sensor_parity(h2, h2_oracle_tables, kind):
    case = reacting_box_case(20, 16, 12, seed=33, tables=h2)
    case.run.update(viscous_order=0, reaction=0, inviscid_scheme=72, shock_sensor=kind, sensor_eps=1e4,
                    sensor_threshold=0.05)
    s, o = pair(case, h2, h2_oracle_tables)
    s.op(_abi.OP_SHOCK_SENSOR)
    o.op("shock_sensor")
    a = s.download(0, _abi.F_SHOCK_SENSOR)[None]
    b = o.a["shock_sensor"][None]
    assert np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1e-300)
    s.close()


def median_threshold(o, case):
    """A threshold that splits the positive sensor values of the interior in
    two halves (Ducros is 0 wherever the flow expands)."""
    o.op("shock_sensor")
    v = interior(case, o.a["shock_sensor"])
    t = float(np.median(v[v > 0]))
    assert t > 1e-10
    return t


@pytest.mark.parametrize("scheme", [72, 71, 52])
@pytest.mark.parametrize("kind", [0, 1])
def test_hybrid_operator_parity(h2, h2_oracle_tables, scheme, kind):
    case = reacting_box_case(20, 16, 12, seed=34, tables=h2)
    case.run.update(viscous_order=0, reaction=0, inviscid_scheme=scheme, shock_sensor=kind, sensor_eps=1e4,
                    sensor_threshold=1.0)
    probe = Oracle(case, h2_oracle_tables)
    probe.initialize()
    case.run.update(sensor_threshold=median_threshold(probe, case))
    s, o = pair(case, h2, h2_oracle_tables)
    s.op(_abi.OP_SHOCK_SENSOR)
    o.op("shock_sensor")
    for d in "xyz":
        s.op(_abi.OP_ZERO_DQ)
        o.op("zero_dq")
        s.op(getattr(_abi, "OP_INVISCID_" + d.upper()))
        o.op("inviscid_" + d)
        check("dq", s.download(0, _abi.F_DQ), o.a["dq"], STAGE_TOL, f" after inviscid_{d}")
    s.close()


@pytest.mark.parametrize("scheme", [71, 72])
def test_hybrid_limited_steps_parity(h2, h2_oracle_tables, scheme):
    case = species_front_case(h2, seed=12)
    case.run.update(inviscid_scheme=scheme, viscous_order=8, reaction=1, shock_sensor=1, sensor_threshold=1e-3,
                    positive_preserving=1, fixed_time_step=1)
    probe = Oracle(case, h2_oracle_tables)
    probe.initialize()
    dt = 0.5 * probe.cfl_dt()
    case.run.update(dt=dt)
    s, o = pair(case, h2, h2_oracle_tables)
    for _ in range(10):
        s.step(dt)
        o.step(dt, fixed=True)
    for name, fid in FIELDS:
        check(name, interior(case, s.download(0, fid)), interior(case, o.a[name]), LONG_TOL, " after 10 steps")
    y = interior(case, s.download(0, _abi.F_SV))
    assert y.min() >= 0.0
    assert np.abs(y.sum(axis=0) - 1.0).max() <= 1e-12
    s.close()
