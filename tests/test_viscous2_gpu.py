"""Parity of the 2nd-order viscous face scheme (src/ViscousScheme.cu:836-1512)
on the GPU with the oracle: the operator alone and full steps, on curvilinear
blocks for the Air perfect gas and the 9-species mixture (with and without
the pressure-gradient diffusion term), and in 2-D."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, channel_case, load_tables, reacting_box_case
from oracle import Oracle

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-11
LONG_TOL = 1e-9


def normwise(a, b, vector=None):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    den = np.abs(b).max(axis=1)
    if vector is not None:
        den[list(vector)] = den[list(vector)].max()
    den = np.maximum(den, 1e-300)
    return np.abs(a - b).max(axis=1) / den


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def make(kind, gradp, tables):
    if kind == "air":
        c = channel_case(24, 16, 12, physics="air", subsonic=True, seed=5)
        c.run.update(viscous_order=2)
    elif kind == "air2d":
        c = channel_case(40, 24, 1, physics="air", subsonic=False, seed=6)
        c.run.update(viscous_order=2)
    else:
        c = reacting_box_case(20, 16, 12, seed=41, tables=tables)
        c.run.update(viscous_order=2, gradp_in_diffusion=gradp)
    return c


@pytest.mark.parametrize("kind,gradp", [("air", 0), ("air2d", 0), ("h2", 0), ("h2", 1)])
def test_viscous2_parity(tables, h2_oracle_tables, kind, gradp):
    case = make(kind, gradp, tables)
    print(f"viscous 2nd order {kind} gradp={gradp}", flush=True)
    mix = case.run["mixture"]
    s = Solver(case.run, tables if mix else None)
    s.add_block(case)
    o = Oracle(case, h2_oracle_tables if mix else None)
    s.initialize()
    o.initialize()
    s.op(_abi.OP_ZERO_DQ)
    o.op("zero_dq")
    s.op(_abi.OP_VISCOUS)
    o.op("viscous")
    err = normwise(s.download(0, _abi.F_DQ), o.a["dq"])
    assert np.all(err[1:] <= STAGE_TOL), f"viscous dq: {err}"
    dt = 0.5 * o.cfl_dt()
    for _ in range(5):
        s.step(dt) if case.run["fixed_time_step"] else s.step()
        o.step(dt, fixed=True) if case.run["fixed_time_step"] else o.step(fixed=False)
    for name, fid in (("cv", _abi.F_CV), ("bv", _abi.F_BV)):
        e = normwise(s.download(0, fid), o.a[name], vector=(1, 2, 3))
        assert np.all(e <= LONG_TOL), f"{name} after 5 steps: {e}"
    s.close()
