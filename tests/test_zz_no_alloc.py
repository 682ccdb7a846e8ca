"""The allocation-free reactor batch (corefl_serve_run_noalloc, through the C
ABI) against the reactor server (corefl_serve_open / submit / progress /
poll) on the same requests: bit-identical results, on two calls in a row,
the second with a smaller batch, plus the same with a failing request (a
step cap of 5) in the batch, and the host-side rejection of a batch beyond
COREFL_NOALLOC_MAX_REACTORS."""
import numpy as np
import pytest

from corefl_amd import ReactorServer, h2_air_mass_fractions, load_tables
from corefl_amd.serve import NOALLOC_MAX_REACTORS, ServeError, run_noalloc

pytestmark = pytest.mark.gpu

P_ATM = 101325.0
STEPS = 128


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def requests(tables, n, seed):
    mw = np.array(tables.mw[:9])
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        scheme = ("rok4e", "point-implicit", "rk3")[i % 3]
        # explicit SSP-RK3 at 1e-10 s and point-implicit at 1e-8 s stay stable on these states
        t_end, dt = {"rok4e": (2e-4, 1e-7), "point-implicit": (2e-5, 1e-8), "rk3": (1e-6, 1e-10)}[scheme]
        q = dict(temperature=float(rng.uniform(1000.0, 1500.0)), pressure=float(rng.uniform(1, 5)) * P_ATM,
                 y=h2_air_mass_fractions(float(rng.uniform(0.5, 1.5)), mw), t_end=t_end, dt=dt, scheme=scheme,
                 adaptive=scheme == "rok4e")
        out.append(q)
    return out


def served(tables, reqs):
    srv = ReactorServer(tables, budget_bytes=8 << 20, steps_per_launch=STEPS)
    tk = [srv.submit(**q) for q in reqs]
    res = srv.run_until_done(tk)
    srv.close()
    return [res[t] for t in tk]


def same(a, b):
    """Bit for bit, a NaN of a failed request matching a NaN."""
    num = ("time", "temperature", "pressure")
    return (all(a[k] == b[k] for k in ("status", "reason", "steps", "rejected"))
            and np.array_equal([a[k] for k in num], [b[k] for k in num], equal_nan=True)
            and np.array_equal(a["y"], b["y"], equal_nan=True))


def test_noalloc_batch_matches_the_server_on_two_calls(tables):
    big = requests(tables, 40, seed=3)
    big[7] = dict(big[7], max_steps=5)  # fails for its step cap
    small = requests(tables, 9, seed=4)
    print("allocation-free batch: 40 requests, then 9", flush=True)
    got_big = run_noalloc(tables, big, STEPS)
    got_small = run_noalloc(tables, small, STEPS)
    ref_big, ref_small = served(tables, big), served(tables, small)
    assert got_big[7]["status"] == -5 and got_big[7]["reason"] == 3
    for g, r in zip(got_big, ref_big):
        assert same(g, r), (g, r)
    for g, r in zip(got_small, ref_small):
        assert same(g, r), (g, r)
    assert all(r["status"] == 0 for i, r in enumerate(got_big) if i != 7), [
        (i, r["status"], r["reason"]) for i, r in enumerate(got_big) if r["status"] != 0]
    # a third call after the smaller one: the stale slots of the first call are not read
    again = run_noalloc(tables, big[:3], STEPS)
    for g, r in zip(again, ref_big[:3]):
        assert same(g, r)


def test_noalloc_rejects_batches_beyond_its_limit(tables):
    reqs = requests(tables, 1, seed=5) * (NOALLOC_MAX_REACTORS + 1)
    with pytest.raises(ServeError, match="-1"):
        run_noalloc(tables, reqs, STEPS)
    assert run_noalloc(tables, [], STEPS) == []
