"""Wu's operator splitting on the GPU (temporal_scheme 4,
src/StrangSplitting.cu:687-865) against the oracle on the same inputs: a
reacting 9-species box with 2nd- and 8th-order viscous terms, each reaction
sub-step method (0 explicit, 1 point-implicit with the full Jacobian,
3 SSP-RK3), three steps of a fixed dt (6e-8 s) that the chemical time scale (about 2e-8 s
in the hot gas) splits into several sub-steps.  Bound: 1e-10 normwise per variable after
three steps (each step is a transport evaluation, a chemistry integration of
a few sub-steps and three RK stages); Y >= 0 and sum Y = 1 to 1e-12.  The
per-operator units (Cn, the reaction sub-step, the dq - Cn update) are
checked on their own at 1e-11."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, load_tables, reacting_box_case
from oracle import Oracle, h2_tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def err_normwise(g, r):
    g = g.reshape(r.shape[0], -1)
    r = r.reshape(r.shape[0], -1)
    return np.abs(g - r).max(axis=1) / np.maximum(np.abs(r).max(axis=1), 1e-300)


def wu_case(tables, viscous, method, chem_kernel=0):
    c = reacting_box_case(16, 12, 10, seed=21, tables=tables, T_range=(1000.0, 2100.0), u_rms=60.0)
    c.run.update(temporal_scheme=4, chem_advance_method=method, viscous_order=viscous, fixed_time_step=1,
                 chem_kernel=chem_kernel)
    return c


# chem_kernel 0: the explicit sub-steps run in the kernel compiled for the mechanism; 1: the table-driven kernel
@pytest.mark.parametrize("viscous,method,chem_kernel", [(8, 0, 0), (8, 0, 1), (8, 1, 0), (2, 3, 0), (0, 0, 0)])
def test_wu_steps_match_the_oracle(tables, viscous, method, chem_kernel):
    c = wu_case(tables, viscous, method, chem_kernel)
    s = Solver(c.run, tables)
    s.add_block(c)
    s.initialize()
    o = Oracle(c, h2_tables())
    o.initialize()
    dt = 6.0e-8
    print(f"Wu splitting viscous {viscous} method {method} chem_kernel {chem_kernel}: 3 steps of {dt}", flush=True)
    for _ in range(3):
        s.step(dt)
        o.step(dt, True)
    for name, fid in (("cv", _abi.F_CV), ("bv", _abi.F_BV), ("sv", _abi.F_SV)):
        g = s.download(0, fid).reshape(o.a[name].shape)
        gi, ri = g[:, 4:-4, 4:-4, 4:-4], o.a[name][:, 4:-4, 4:-4, 4:-4]
        err = err_normwise(gi, ri)
        assert np.all(err <= 1e-10), (name, err)
    y = s.download(0, _abi.F_SV).reshape(o.a["sv"].shape)[:, 4:-4, 4:-4, 4:-4]
    assert y.min() >= 0.0 and np.abs(y.sum(axis=0) - 1.0).max() <= 1e-12
    # the chemistry really was split into several sub-steps somewhere
    rts = s.download(0, _abi.F_RTS).reshape(o.a["rts"].shape)[4:-4, 4:-4, 4:-4]
    assert (rts < dt / 2).any()
    s.close()


def test_wu_operators_match_the_oracle(tables):
    c = wu_case(tables, 8, 1)
    s = Solver(c.run, tables)
    s.add_block(c)
    s.initialize()
    o = Oracle(c, h2_tables())
    o.initialize()
    dt = 6.0e-8
    s.step(dt)
    o.step(dt, True)  # a step first, so the reaction time scale is set
    # Cn: zero dq, viscous, inviscid, copy
    for op, name in ((_abi.OP_ZERO_DQ, "zero_dq"), (_abi.OP_VISCOUS, "viscous"), (_abi.OP_INVISCID_X, "inviscid_x"),
                     (_abi.OP_INVISCID_Y, "inviscid_y"), (_abi.OP_INVISCID_Z, "inviscid_z")):
        s.op(op)
        o.op(name)
    s.op(_abi.OP_WU_CN)
    o.a["cn"][:] = o.a["dq"]
    s.op(_abi.OP_WU_REACTION, dt=dt)
    o.lib.oc_wu_reaction(o.args[0], o.args[1], o.args[2], dt)
    for name, fid in (("cv", _abi.F_CV), ("bv", _abi.F_BV), ("sv", _abi.F_SV), ("rts", _abi.F_RTS)):
        g = s.download(0, fid).reshape(o.a[name].shape)
        gi = g[..., 4:-4, 4:-4, 4:-4]
        ri = o.a[name][..., 4:-4, 4:-4, 4:-4]
        if name == "rts":
            gi, ri = gi[None], ri[None]
        assert np.all(err_normwise(gi, ri) <= 1e-11), name
    s.op(_abi.OP_COPY_QN)
    o.op("copy_qn")
    s.op(_abi.OP_UPDATE_RK_WU, rk=0, dt=0.5 * dt)
    o.lib.oc_update_rk_wu(o.args[0], o.args[1], o.args[2], 0.5 * dt, 0)
    for name, fid in (("cv", _abi.F_CV), ("bv", _abi.F_BV)):
        g = s.download(0, fid).reshape(o.a[name].shape)
        assert np.all(err_normwise(g[:, 4:-4, 4:-4, 4:-4], o.a[name][:, 4:-4, 4:-4, 4:-4]) <= 1e-11), name
    s.close()


def test_wu_tubes_match_the_oracle_alone_and_in_a_mix(tables, h2_oracle_tables):
    """Served shock tubes with Wu's splitting at a fixed dt (the reference's reactiveShockTube scheme,
    example/reactiveShockTube/input/setup.txt:19-22): a tube matches the oracle's one-row-block run of
    the same tube (1e-9 normwise after the fixed steps, the last cut to t_end), and tubes of several
    sizes give the same results bit for bit alone and together."""
    from test_tube_server_gpu import march, row, served, states
    from corefl_amd import TubeServer, tube_case
    dt = 1e-8
    run = dict(temporal_scheme=4, fixed_time_step=1, dt=dt)
    left, right = states(tables, T_l=1600.0, p_l=5 * 101325.0, T_r=1000.0, p_r=2 * 101325.0, u_r=-50.0)
    n, length, t_end = 48, 0.02, 2.05e-7
    srv = TubeServer(tables, run=run, budget_bytes=32 << 20, steps_per_progress=4)
    print(f"Wu tube oracle parity n {n}", flush=True)
    t = srv.submit(n, length, t_end, left, right, n_left=20, bc=("symmetry", "outflow"))
    res = served(srv, [t], "wu tube")[t]
    srv.close()
    case = tube_case(n, length, left, right, tables, n_left=20, bc=("symmetry", "outflow"))
    case.run.update(temporal_scheme=4, fixed_time_step=1, dt=dt)
    o = Oracle(case, h2_oracle_tables)
    o.initialize()
    steps = march(o, t_end, lambda: dt, lambda d: o.step(d, fixed=True))
    assert res["status"] == 0 and res["steps"] == steps == 21 and res["time"] == t_end
    ref = np.concatenate([row(case, o.a["bv"]), row(case, o.a["sv"])])
    scale = np.abs(ref).max(axis=1)
    scale[1:4] = scale[1:4].max()
    err = np.abs(res["profile"] - ref).max(axis=1) / np.maximum(scale, 1e-300)
    assert np.all(err <= 1e-9), err
    # alone and in a mix of sizes
    mw = np.array(tables.mw[:9])
    reqs = [dict(n_cells=m, length=0.0004 * m, t_end=1.5e-7, left=left, right=right, bc=("symmetry", "symmetry"))
            for m in (24, 64, 40)]
    alone = []
    for q in reqs:
        srv = TubeServer(tables, run=run, budget_bytes=16 << 20, steps_per_progress=4)
        tk = srv.submit(**q)
        alone.append(served(srv, [tk], "wu alone")[tk])
        srv.close()
    srv = TubeServer(tables, run=run, budget_bytes=32 << 20, steps_per_progress=4)
    tks = [srv.submit(**q) for q in reqs]
    mix = served(srv, tks, "wu mix")
    srv.close()
    for a, tk in zip(alone, tks):
        assert a["steps"] == mix[tk]["steps"] and np.array_equal(a["profile"], mix[tk]["profile"])
