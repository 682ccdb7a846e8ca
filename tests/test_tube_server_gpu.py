"""The shock-tube request server on the GPU.  A served tube is the one-block
run of the same tube: every row of a (n, 2 g + 1, 1) block holds the tube and
the solver takes the same CFL steps (each step's dt from local_time_step, the
last one cut to t_end).  The block's y-boundary kernels rebuild the ghost
rows' conservative variables from the primitives (outflow and periodic alike),
which differ from the updated ones in the last bits, so the block's rows part
at rounding level and the middle row matches the tube to 1e-9 (velocities
normalised as a vector), as it matches the oracle's run of the block; the tube
itself keeps v exactly 0.  Bit for bit: each result is the same alone and
among tubes of other sizes arriving in two waves, and a server whose budget
holds two tubes at a time (so tubes reuse row space) queues the rest and
returns the same results.  Hot reacting tubes take steps of a few ns (local_time_step's
chemical time-scale bound), so the tubes run for a microsecond or less."""
import sys

import numpy as np
import pytest

from corefl_amd import ServeError, Solver, TubeServer, _abi, h2_air_mass_fractions, load_tables, tube_case
from oracle import Oracle

pytestmark = pytest.mark.gpu

P_ATM = 101325.0
LONG_TOL = 1e-9


@pytest.fixture(scope="module")
def h2():
    return load_tables()


def states(tables, T_l=1800.0, p_l=8 * P_ATM, T_r=950.0, p_r=1 * P_ATM, phi_l=1.0, phi_r=1.0, u_l=0.0, u_r=0.0):
    mw = np.array(tables.mw[:9])
    return (T_l, p_l, u_l, h2_air_mass_fractions(phi_l, mw)), (T_r, p_r, u_r, h2_air_mass_fractions(phi_r, mw))


def served(srv, tickets, label):
    """run_until_done with a line per 20 progress calls."""
    out, calls = {}, 0
    while True:
        pending = srv.progress()
        calls += 1
        if calls % 20 == 0:
            print(f"{label}: {calls} progress calls, {pending} pending", file=sys.stderr, flush=True)
        for t in srv.ready():
            if t in tickets:
                out[t] = srv.poll(t)
        if pending == 0:
            return out


def march(sim, t_end, op_dt, op_step):
    """The server's clock: CFL step from the current state, cut to t_end."""
    time, steps = 0.0, 0
    while True:
        dt = op_dt()
        last = dt >= t_end - time
        if last:
            dt = t_end - time
        op_step(dt)
        steps += 1
        if last:
            return steps
        time += dt


def err_normwise(a, ref):
    scale = np.abs(ref).max(axis=1)
    scale[1:4] = scale[1:4].max()
    return np.abs(a - ref).max(axis=1) / np.maximum(scale, 1e-300)


def row(case, arr):
    g = case.ngg
    return arr[:, g, 2 * g, g:g + case.mx]


@pytest.mark.parametrize("scheme,bc", [(72, ("outflow", "symmetry")), (52, ("symmetry", "outflow"))])
def test_tube_matches_one_block_run(h2, scheme, bc):
    left, right = states(h2)
    n, length, t_end = 96, 0.05, 1e-6
    srv = TubeServer(h2, run=dict(inviscid_scheme=scheme), budget_bytes=64 << 20, steps_per_progress=8)
    print(f"tube scheme {scheme} bc {bc} n {n}", file=sys.stderr, flush=True)
    t = srv.submit(n, length, t_end, left, right, n_left=40, bc=bc)
    res = served(srv, [t], "tube")[t]
    print(f"served in {res['steps']} steps", file=sys.stderr, flush=True)
    srv.close()
    case = tube_case(n, length, left, right, h2, n_left=40, bc=bc, scheme=scheme)
    case.run.update(fixed_time_step=1)
    s = Solver(case.run, h2)
    s.add_block(case)
    s.initialize()
    steps = march(s, t_end, s.cfl_dt, s.step)
    assert res["steps"] == steps and res["time"] == t_end
    ref = np.concatenate([row(case, s.download(0, _abi.F_BV)), row(case, s.download(0, _abi.F_SV))])
    s.close()
    assert np.all(err_normwise(res["profile"], ref) <= LONG_TOL), err_normwise(res["profile"], ref)
    assert np.all(res["profile"][2] == 0.0) and np.all(res["profile"][3] == 0.0)
    # the flow moved: the shock left the diaphragm and the reflecting end heated the gas
    assert np.ptp(res["profile"][1]) > 1.0


def test_tube_oracle_parity(h2, h2_oracle_tables):
    left, right = states(h2, T_l=1600.0, p_l=5 * P_ATM, T_r=1000.0, p_r=2 * P_ATM, u_r=-50.0)
    n, length, t_end = 48, 0.02, 5e-7
    srv = TubeServer(h2, budget_bytes=32 << 20, steps_per_progress=4)
    print(f"tube oracle parity n {n}", file=sys.stderr, flush=True)
    t = srv.submit(n, length, t_end, left, right, n_left=20, bc=("symmetry", "outflow"))
    res = served(srv, [t], "tube")[t]
    srv.close()
    case = tube_case(n, length, left, right, h2, n_left=20, bc=("symmetry", "outflow"))
    o = Oracle(case, h2_oracle_tables)
    o.initialize()
    steps = march(o, t_end, o.cfl_dt, lambda dt: o.step(dt, fixed=True))
    assert res["steps"] == steps
    ref = np.concatenate([row(case, o.a["bv"]), row(case, o.a["sv"])])
    err = err_normwise(res["profile"], ref)
    assert np.all(err <= LONG_TOL), f"normwise error per variable {err}"
    y = res["profile"][6:]
    assert y.min() >= 0.0 and np.abs(y.sum(axis=0) - 1.0).max() <= 1e-12


def tube_requests(tables, seed):
    rng = np.random.default_rng(seed)
    mw = np.array(tables.mw[:9])
    out = []
    for n in (16, 40, 64, 100, 130, 200, 72):
        T_l, T_r = rng.uniform(1200, 2000), rng.uniform(900, 1200)
        p_l, p_r = rng.uniform(3, 10) * P_ATM, rng.uniform(1, 3) * P_ATM
        y_l, y_r = (h2_air_mass_fractions(rng.uniform(0.5, 2.0), mw) for _ in range(2))
        bc = ("outflow", "symmetry") if rng.uniform() < 0.5 else ("symmetry", "symmetry")
        out.append(dict(n_cells=n, length=0.001 * n, t_end=float(rng.uniform(2e-7, 8e-7)),
                        left=(T_l, p_l, 0.0, y_l), right=(T_r, p_r, 0.0, y_r),
                        n_left=int(rng.integers(4, n - 4)), bc=bc))
    return out


def test_alone_and_in_the_mix_are_identical(h2):
    reqs = tube_requests(h2, seed=3)
    alone = []
    for q in reqs:
        srv = TubeServer(h2, budget_bytes=16 << 20, steps_per_progress=8)
        t = srv.submit(**q)
        alone.append(served(srv, [t], f"alone {q['n_cells']}")[t])
        srv.close()
    srv = TubeServer(h2, budget_bytes=64 << 20, steps_per_progress=8)
    print("tubes in two waves", file=sys.stderr, flush=True)
    tickets = [srv.submit(**q) for q in reqs[:4]]
    srv.progress()
    tickets += [srv.submit(**q) for q in reqs[4:]]
    res = served(srv, tickets, "mix")
    srv.close()
    for a, t in zip(alone, tickets):
        b = res[t]
        assert a["steps"] == b["steps"] and a["time"] == b["time"]
        assert np.array_equal(a["profile"], b["profile"])


def test_budget_admission(h2):
    reqs = tube_requests(h2, seed=4)
    big = TubeServer(h2, budget_bytes=64 << 20, steps_per_progress=8)
    # room for the two largest tubes at a time
    small = TubeServer(h2, budget_bytes=int(2.3 * 208 * 700), steps_per_progress=8)
    assert small.capacity < 2 * (200 + 8) + 200
    tb = [big.submit(**q) for q in reqs]
    ts = [small.submit(**q) for q in reqs]
    rb, rs = served(big, tb, "big"), served(small, ts, "small")
    for a, b in zip(tb, ts):
        assert np.array_equal(rb[a]["profile"], rs[b]["profile"])
    big.close()
    small.close()


def test_bad_requests(h2):
    srv = TubeServer(h2, budget_bytes=1 << 20)
    left, right = states(h2)
    with pytest.raises(ServeError):
        srv.submit(8, 0.01, 1e-5, left, right)  # fewer than 16 nodes
    with pytest.raises(ServeError):
        srv.submit(10 ** 6, 1.0, 1e-5, left, right)  # longer than the row
    with pytest.raises(ServeError):
        srv.submit(32, 0.01, 1e-5, left, right, bc=(5, "outflow"))  # inflow end
    srv.close()
    with pytest.raises(ServeError):
        TubeServer(h2, run=dict(viscous_order=2))
