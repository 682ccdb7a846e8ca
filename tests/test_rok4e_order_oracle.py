"""The oracle's ROK4E reactor (src/ConstVolumeReactor.cpp:442-707, tableau
src/Define.h:33-48) pinned by its order of accuracy with fixed steps
(adaptive = 0):

* on a stiff linear chain A -> B -> C (first-order irreversible reactions with
  temperature-independent rates and three species of identical thermodynamics,
  so T stays put and the source Jacobian is exact) the error against the
  closed-form solution falls by 2^4 per halving of the step: fourth order;
* on the nonlinear kinetics of the H2 mechanism (its 2 <-> 2 reactions with
  temperature-independent thermodynamics, so the Jacobian the reference
  forms is exact) the error against a 64x finer run falls by 2^4 per halving
  as well;
* the explicit SSP-RK3 reactor shows order 3 on the same chain, so the
  measurement tells the orders apart.

A lower-order or mis-staged scheme fails these (an equilibrium end state or
RTol self-convergence does not tell).  With the full mechanism the
reference's Jacobian leaves out the temperature and third-body dependences,
so ROK4E runs as a W-method there and its observed order is lower (1-2 near
equilibrium, measured on the oracle); the GPU reactors follow the oracle's
arithmetic either way."""
import ctypes as C

import numpy as np

from oracle import h2_tables, load

K1, K2 = 400.0, 40.0  # 1/s: stiffness ratio 10, h K1 from 0.4 down to 0.05
RHO = 1.0
T0 = 1000.0
P_ATM = 101325.0


def chain_tables():
    t = h2_tables()
    n2 = 8
    low, high = list(t.low_coeff[n2]), list(t.high_coeff[n2])
    mw = t.mw[n2]
    t.n_spec, t.n_reac = 3, 2
    for s in range(3):
        t.mw[s], t.imw[s], t.gas_const[s] = mw, 1.0 / mw, 8314.462618 / mw
        t.t_low[s], t.t_mid[s], t.t_high[s] = t.t_low[n2], t.t_mid[n2], t.t_high[n2]
        t.low_coeff[s][:] = low
        t.high_coeff[s][:] = high
    for r, (k, f, b) in enumerate(((K1, 0, 1), (K2, 1, 2))):
        t.reac_type[r], t.rev_type[r], t.order[r] = 0, 0, 0
        t.A[r], t.b[r], t.Ea[r] = k, 0.0, 0.0
        for s in range(20):
            t.stoi_f[r][s] = 1 if s == f else 0
            t.stoi_b[r][s] = 1 if s == b else 0
            t.third_body[r][s] = 0.0
    return t


def chain_exact(y0, t):
    a0, b0, c0 = y0
    a = a0 * np.exp(-K1 * t)
    b = b0 * np.exp(-K2 * t) + a0 * K1 / (K2 - K1) * (np.exp(-K1 * t) - np.exp(-K2 * t))
    return np.array([a, b, a0 + b0 + c0 - a - b])


def run_rok4e(tb, y0, T, h, t_end):
    y = np.ascontiguousarray(y0, float).copy()
    t = C.c_double(T)
    rej = C.c_long(0)
    steps = load().oc_cv_reactor_rok4e(C.byref(tb), RHO, C.byref(t), y.ctypes.data_as(C.POINTER(C.c_double)), h,
                                        t_end, 0, 1e-4, 1e-8, C.byref(rej))
    assert rej.value == 0
    return y, t.value, steps


def run_rk3(tb, y0, T, h, t_end):
    y = np.ascontiguousarray(y0, float).copy()
    t = C.c_double(T)
    load().oc_cv_reactor_ex(C.byref(tb), RHO, C.byref(t), y.ctypes.data_as(C.POINTER(C.c_double)), h, t_end, 3, 0)
    return y, t.value


def test_rok4e_fourth_order_on_a_stiff_linear_chain():
    tb = chain_tables()
    y0 = [0.6, 0.3, 0.1]  # every species present: the reference Jacobian drops columns of absent species
    t_end = 2.0 ** -6  # steps of 2^-k s add up exactly
    errs = []
    for n in (16, 32, 64, 128):
        y, T, steps = run_rok4e(tb, y0, T0, t_end / n, t_end)
        assert steps == n and abs(T - T0) <= 1e-6 * T0  # the Newton on E stops at 1e-8 relative steps
        errs.append(np.abs(y - chain_exact(y0, t_end)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    # 3.65, 3.80, 3.90: tending to 4 as h K1 falls (a third-order scheme stays at 3)
    assert np.all(rates > 3.6) and rates[-1] > 3.85 and np.all(np.diff(rates) > 0) and np.all(rates < 4.3), \
        (errs, rates)


def test_rk3_third_order_on_the_same_chain():
    tb = chain_tables()
    y0 = [0.6, 0.3, 0.1]
    t_end = 2.0 ** -6  # steps of 2^-k s add up exactly
    errs = []
    for n in (64, 128, 256):
        y, _ = run_rk3(tb, y0, T0, t_end / n, t_end)
        errs.append(np.abs(y - chain_exact(y0, t_end)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 2.8) and np.all(rates < 3.3), (errs, rates)


def h2_isothermal_tables():
    """Li-2004 with every species given N2's NASA polynomials and the
    third-body / fall-off reactions switched off (A = A2 = 0): what remains
    are the mechanism's 2 <-> 2 reactions (13 of 19, duplicates included),
    which keep the mole number, so T and every rate constant stay put and the
    Jacobian the reference builds (mass action only) is exact."""
    tb = h2_tables()
    n2 = 8
    for s in range(tb.n_spec):
        tb.low_coeff[s][:] = list(tb.low_coeff[n2])
        tb.high_coeff[s][:] = list(tb.high_coeff[n2])
        tb.t_low[s], tb.t_mid[s], tb.t_high[s] = tb.t_low[n2], tb.t_mid[n2], tb.t_high[n2]
    off = 0
    for r in range(tb.n_reac):
        if tb.reac_type[r] >= 4:
            tb.A[r] = tb.A2[r] = 0.0
            off += 1
    assert off == 6
    return tb


def test_rok4e_fourth_order_on_h2_kinetics():
    global RHO
    tb = h2_isothermal_tables()
    ns = tb.n_spec
    # a partly burnt H2/air mixture: every species present, radicals far from equilibrium
    y0 = np.array([0.02, 0.001, 0.2, 0.002, 0.01, 1e-4, 1e-5, 0.05, 0.0])
    y0[8] = 1.0 - y0.sum()
    imw = (y0 / np.array(tb.mw[:ns])).sum()
    rho_keep = RHO
    RHO = P_ATM / (8314.462618 * imw * 1500.0)
    try:
        t_end = 2.0 ** -24  # 60 ns: the radicals move by O(1e-2) in mass fraction
        ref, T_ref, _ = run_rok4e(tb, y0, 1500.0, t_end / 16384, t_end)
        assert abs(T_ref - 1500.0) <= 1e-9 * 1500.0
        assert np.abs(ref - y0).max() > 1e-3
        errs = []
        for n in (32, 64, 128, 256):
            y, _, _ = run_rok4e(tb, y0, 1500.0, t_end / n, t_end)
            errs.append(np.abs(y - ref).max())
        rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
        assert np.all(rates > 3.75) and np.all(rates < 4.3), (errs, rates)
    finally:
        RHO = rho_keep
