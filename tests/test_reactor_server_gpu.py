"""The reactor request server on the GPU: every scheme against the oracle's
reactor on the same request (fixed steps to 1e-9, adaptive ROK4E to its
tolerance), each result identical bit for bit alone and in a mix of two
hundred concurrent requests (and of 2,100, which the server steps with the
one-lane-per-slot kernel until few remain), and admission by slot budget (a three-slot
server queues the rest and returns the same results).  The explicit SSP-RK3
runs at dt = 2e-10 s: at 1e-9 s it sits at its stability limit through the
ignition of the 1400 K case, where the oracle itself moves by 4e-6 for a
1e-15 change of the initial temperature."""
import ctypes as C

import numpy as np
import pytest

from corefl_amd import ReactorServer, h2_air_mass_fractions, load_tables
from oracle import PD, h2_tables, load

pytestmark = pytest.mark.gpu

R_U = 8314.462618
P_ATM = 101325.0


@pytest.fixture(scope="module")
def tables():
    return load_tables()


@pytest.fixture(scope="module")
def oc():
    return h2_tables()


def oracle_reactor(oc, T0, p0, y, scheme, dt, t_end, adaptive=1, rtol=1e-4, atol=1e-8):
    mw = np.array(oc.mw[:9])
    rho = p0 / (R_U * (y / mw).sum() * T0)
    T = C.c_double(T0)
    yy = np.ascontiguousarray(np.concatenate([y, np.zeros(11)]))
    L = load()
    if scheme == 1:
        rej = C.c_long(0)
        n = L.oc_cv_reactor_rok4e(C.byref(oc), rho, C.byref(T), yy.ctypes.data_as(PD), dt, t_end, adaptive, rtol,
                                  atol, C.byref(rej))
    else:
        n = L.oc_cv_reactor_ex(C.byref(oc), rho, C.byref(T), yy.ctypes.data_as(PD), dt, t_end, scheme, 1)
    return T.value, yy[:9].copy(), n


CASES = [(1100.0, 2 * P_ATM, 1.0), (1400.0, 5 * P_ATM, 0.6), (950.0, 1 * P_ATM, 1.8)]


@pytest.mark.parametrize("scheme,dt,t_end", [(1, 2e-8, 1e-4), (0, 1e-8, 1e-4), (3, 2e-10, 2e-6)])
def test_fixed_step_parity(tables, oc, scheme, dt, t_end):
    mw = np.array(tables.mw[:9])
    srv = ReactorServer(tables, budget_bytes=1 << 20, steps_per_launch=4096)
    tickets = []
    for T0, p0, phi in CASES:
        y = h2_air_mass_fractions(phi, mw)
        tickets.append(srv.submit(T0, p0, y, t_end, dt, scheme=scheme, adaptive=False, conserve_mass=True))
    print(f"scheme {scheme} dt {dt} t_end {t_end}", flush=True)
    res = srv.run_until_done(tickets)
    for (T0, p0, phi), t in zip(CASES, tickets):
        r = res[t]
        To, yo, n = oracle_reactor(oc, T0, p0, h2_air_mass_fractions(phi, mw), scheme, dt, t_end, adaptive=0)
        assert r["steps"] == n
        assert abs(r["temperature"] - To) <= 1e-9 * To, (r["temperature"], To)
        assert np.abs(r["y"] - yo).max() <= 1e-9 * np.abs(yo).max()
    srv.close()


def test_adaptive_rok4e_parity(tables, oc):
    mw = np.array(tables.mw[:9])
    srv = ReactorServer(tables, budget_bytes=1 << 20, steps_per_launch=512)
    tickets = [srv.submit(T0, p0, h2_air_mass_fractions(phi, mw), 2e-3, 1e-8, rtol=1e-6, atol=1e-10)
               for T0, p0, phi in CASES]
    res = srv.run_until_done(tickets)
    for (T0, p0, phi), t in zip(CASES, tickets):
        r = res[t]
        To, yo, n = oracle_reactor(oc, T0, p0, h2_air_mass_fractions(phi, mw), 1, 1e-8, 2e-3, rtol=1e-6, atol=1e-10)
        # the step sequences may part where an error estimate sits at 1 to rounding;
        # both integrate to the tolerance
        assert abs(r["temperature"] - To) <= 1e-5 * To, (r["temperature"], To)
        assert np.abs(r["y"] - yo).max() <= 1e-4 * np.abs(yo).max()
        assert abs(r["steps"] - n) <= 0.05 * n
    srv.close()


def random_requests(mw, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        T0, p0, phi = rng.uniform(900, 1600), rng.uniform(1, 10) * P_ATM, rng.uniform(0.5, 2.0)
        out.append((T0, p0, h2_air_mass_fractions(phi, mw)))
    return out


def test_alone_and_in_the_mix_are_identical(tables):
    mw = np.array(tables.mw[:9])
    reqs = random_requests(mw, 200, seed=5)
    pick = [0, 57, 199]
    alone = {}
    for i in pick:
        srv = ReactorServer(tables, budget_bytes=1 << 20, steps_per_launch=256)
        T0, p0, y = reqs[i]
        t = srv.submit(T0, p0, y, 5e-4, 1e-8)
        alone[i] = srv.run_until_done([t])[t]
        srv.close()
    srv = ReactorServer(tables, budget_bytes=8 << 20, steps_per_launch=256)
    # requests arrive in two waves: the second joins while the first is in flight
    tickets = [srv.submit(T0, p0, y, 5e-4, 1e-8) for T0, p0, y in reqs[:100]]
    srv.progress()
    tickets += [srv.submit(T0, p0, y, 5e-4, 1e-8) for T0, p0, y in reqs[100:]]
    res = srv.run_until_done(tickets)
    assert len(res) == 200
    for i in pick:
        r = res[tickets[i]]
        assert r["temperature"] == alone[i]["temperature"]
        assert np.array_equal(r["y"], alone[i]["y"])
        assert r["steps"] == alone[i]["steps"] and r["rejected"] == alone[i]["rejected"]
    srv.close()


def test_budget_admission(tables):
    mw = np.array(tables.mw[:9])
    reqs = random_requests(mw, 20, seed=6)
    small = ReactorServer(tables, budget_bytes=4 * 400, steps_per_launch=128)
    assert 1 <= small.capacity <= 8
    big = ReactorServer(tables, budget_bytes=8 << 20, steps_per_launch=128)
    ts = [small.submit(T0, p0, y, 2e-4, 1e-8) for T0, p0, y in reqs]
    tb = [big.submit(T0, p0, y, 2e-4, 1e-8) for T0, p0, y in reqs]
    rs, rb = small.run_until_done(ts), big.run_until_done(tb)
    for a, b in zip(ts, tb):
        assert rs[a]["temperature"] == rb[b]["temperature"]
        assert np.array_equal(rs[a]["y"], rb[b]["y"])
    small.close()
    big.close()


def test_result_does_not_depend_on_the_kernel_the_load_picks(tables):
    """With more than 2,048 slots in flight the server steps them one lane per
    slot, with fewer one wave per slot: 2,100 requests start on the first and
    finish on the second, and the requests picked come back bit for bit as
    they do alone (one wave per slot from start to end)."""
    mw = np.array(tables.mw[:9])
    reqs = random_requests(mw, 2100, seed=8)
    pick = [0, 1033, 2099]
    alone = {}
    for i in pick:
        srv = ReactorServer(tables, budget_bytes=1 << 20, steps_per_launch=64)
        T0, p0, y = reqs[i]
        t = srv.submit(T0, p0, y, 1e-4, 1e-8)
        alone[i] = srv.run_until_done([t])[t]
        srv.close()
    srv = ReactorServer(tables, budget_bytes=16 << 20, steps_per_launch=64)
    print("2,100 reactors in one server", flush=True)
    tickets = [srv.submit(T0, p0, y, 1e-4, 1e-8) for T0, p0, y in reqs]
    res = srv.run_until_done(tickets)
    srv.close()
    assert len(res) == 2100 and all(r["status"] == 0 for r in res.values())
    for i in pick:
        r = res[tickets[i]]
        assert r["temperature"] == alone[i]["temperature"] and np.array_equal(r["y"], alone[i]["y"])
        assert r["steps"] == alone[i]["steps"] and r["rejected"] == alone[i]["rejected"]
