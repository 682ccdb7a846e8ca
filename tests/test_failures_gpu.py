"""Failures report instead of hanging (the reference stops on a non-finite
residual, src/Residual.cu:97-118).  Every condition is forced from the host
and every kernel exits normally:

* a served reactor with a tiny step cap fails alone (status COREFL_ERR_STATE,
  reason COREFL_FAIL_CAP) and frees its slot, while the requests batched with
  it come back bit for bit as they do alone;
* the same for a served shock tube with a step cap;
* a 3-D state uploaded with a negative density at one interior node makes
  corefl_sync and the next corefl_step / corefl_stage return
  COREFL_ERR_STATE (the update kernel flags the point), and uploading a good
  state clears the flag."""
import numpy as np
import pytest

from corefl_amd import ReactorServer, Solver, TubeServer, _abi, h2_air_mass_fractions, load_tables, reacting_box_case
from corefl_amd.solver import CoreflError

pytestmark = pytest.mark.gpu

P_ATM = 101325.0
ERR_STATE, FAIL_CAP = -5, 3


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def test_reactor_step_cap_fails_alone(tables):
    mw = np.array(tables.mw[:9])
    reqs = [(1100.0, 2 * P_ATM, 1.0), (1400.0, 5 * P_ATM, 0.6), (950.0, 1 * P_ATM, 1.8)]
    alone = []
    for T0, p0, phi in reqs:
        srv = ReactorServer(tables, budget_bytes=1 << 20, steps_per_launch=256)
        t = srv.submit(T0, p0, h2_air_mass_fractions(phi, mw), 5e-4, 1e-8)
        alone.append(srv.run_until_done([t])[t])
        srv.close()
    srv = ReactorServer(tables, budget_bytes=1 << 20, steps_per_launch=256)
    tickets = [srv.submit(T0, p0, h2_air_mass_fractions(phi, mw), 5e-4, 1e-8) for T0, p0, phi in reqs[:2]]
    bad = srv.submit(1200.0, 3 * P_ATM, h2_air_mass_fractions(1.0, mw), 5e-4, 1e-8, max_steps=5)
    tickets.append(srv.submit(*reqs[2][:2], h2_air_mass_fractions(reqs[2][2], mw), 5e-4, 1e-8))
    print("reactors: three requests and one with a step cap of 5", flush=True)
    res = srv.run_until_done(tickets + [bad])
    srv.close()
    rb = res[bad]
    assert rb["status"] == ERR_STATE and rb["reason"] == FAIL_CAP
    assert rb["steps"] + rb["rejected"] == 5 and rb["time"] < 5e-4
    for a, t in zip(alone, tickets):
        b = res[t]
        assert b["status"] == 0 and b["time"] == 5e-4
        assert a["temperature"] == b["temperature"] and a["steps"] == b["steps"]
        assert np.array_equal(a["y"], b["y"])


def test_tube_step_cap_fails_alone(tables):
    mw = np.array(tables.mw[:9])
    left = (1800.0, 8 * P_ATM, 0.0, h2_air_mass_fractions(1.0, mw))
    right = (950.0, 1 * P_ATM, 0.0, h2_air_mass_fractions(1.0, mw))
    # 2 us: past the first CFL step (5e-7 s at 1 mm), the chemical time scale bounds the steps
    reqs = [dict(n_cells=n, length=0.001 * n, t_end=2e-6, left=left, right=right) for n in (40, 72, 100)]
    alone = []
    for q in reqs:
        srv = TubeServer(tables, budget_bytes=16 << 20, steps_per_progress=8)
        t = srv.submit(**q)
        alone.append(srv.run_until_done([t])[t])
        srv.close()
    srv = TubeServer(tables, budget_bytes=64 << 20, steps_per_progress=8)
    tickets = [srv.submit(**reqs[0])]
    bad = srv.submit(n_cells=64, length=0.064, t_end=2e-6, left=left, right=right, max_steps=3)
    tickets += [srv.submit(**q) for q in reqs[1:]]
    print("tubes: three requests and one with a step cap of 3", flush=True)
    res = srv.run_until_done(tickets + [bad])
    srv.close()
    rb = res[bad]
    assert rb["status"] == ERR_STATE and rb["reason"] == FAIL_CAP and rb["steps"] == 3 and rb["time"] < 2e-6
    for a, t in zip(alone, tickets):
        b = res[t]
        assert b["status"] == 0 and a["steps"] == b["steps"] > 3 and a["time"] == b["time"]
        assert np.array_equal(a["profile"], b["profile"])


def test_block_state_flag(tables):
    case = reacting_box_case(16, 12, 12, seed=7, tables=tables)
    good_bv, good_sv = case.bv.copy(), case.sv.copy()
    g = case.ngg
    case.bv[0, g + 5, g + 6, g + 7] = -case.bv[0, g + 5, g + 6, g + 7]  # one interior node
    s = Solver(case.run, tables)
    s.add_block(case)
    s.initialize()
    dt = 1e-9
    print("3-D block with a negative density at one node: one step", flush=True)
    s.step(dt)  # the flag is posted behind the step; this call does not wait for it
    with pytest.raises(CoreflError) as e:
        s.sync()
    assert e.value.status == ERR_STATE
    with pytest.raises(CoreflError) as e:
        s.step(dt)
    assert e.value.status == ERR_STATE
    # a good state clears the flag
    s.upload(0, good_bv, good_sv)
    s.initialize()
    s.step(dt)
    s.sync()
    s.close()
