"""The oracle's two-temperature operators pinned by reductions and invariants:

* with no vibrational or electronic modes (theta_v = 0, no levels) and the
  reaction temperature T, the two-temperature path reduces to the
  one-temperature path: E_ve stays 0 and ten steps agree with the 1T oracle
  (2nd- and 8th-order viscous terms);
* V-T relaxation moves energy between the modes only (rho E unchanged), and
  over a step much longer than tau it lands on E_ve = E_el(Tve) + E_v(T)
  (the exact exponential with alpha = 0), src/FieldOperation.cuh:60-88;
* in a still gas of uniform T and composition, the viscous energy flux is
  the kappa_ve grad Tve conduction alone and equals the E_ve flux, so the
  energy and E_ve rows of the viscous divergence coincide (2nd- and
  8th-order schemes);
* with reaction-temperature exponents a = 0, b = 1 the two-temperature
  chemistry at (T, Tve) equals the one-temperature chemistry at the smoothed
  Tve, src/FiniteRateChem.cu:1490-1589."""
import ctypes as C
import math

import numpy as np
import pytest

from corefl_amd import air5_box_case, load_air5_tables
from oracle import PD, Oracle, air5_tables, load


def no_modes(t):
    for i in range(t.n_spec):
        t.theta_v[i] = 0.0
        t.n_elec[i] = 0
    t.trx_model = 0
    return t


def interior(c, a):
    g = c.ngg
    return a[..., g:-g, g:-g, g:-g]


@pytest.mark.parametrize("order", [2, 8])
def test_two_temperature_reduces_to_one_temperature(order):
    t2 = no_modes(air5_tables(two_temperature=True))
    t1 = air5_tables(two_temperature=False)
    g = 4 if order == 8 else 3
    c2 = air5_box_case(14, 12, 10, seed=8, tables=load_air5_tables(False), two_temperature=True, ngg=g)
    c1 = air5_box_case(14, 12, 10, seed=8, tables=load_air5_tables(False), two_temperature=False, ngg=g)
    c1.run.update(viscous_order=order)
    c2.run.update(viscous_order=order)
    o2, o1 = Oracle(c2, t2), Oracle(c1, t1)
    o2.initialize()
    o1.initialize()
    assert np.abs(o2.a["sv"][5]).max() == 0.0
    dt = 0.5 * o1.cfl_dt()
    for _ in range(10):
        o1.step(dt, fixed=True)
        o2.step(dt, fixed=True)
    for name in ("bv", "cv"):
        a1 = interior(c1, o1.a[name])
        a2 = interior(c2, o2.a[name])[:a1.shape[0]]
        err = np.abs(a2 - a1).reshape(a1.shape[0], -1).max(axis=1) / np.abs(a1).reshape(a1.shape[0], -1).max(axis=1)
        assert np.all(err < 1e-10), (name, err)
    assert np.abs(interior(c2, o2.a["cv"])[10]).max() == 0.0


def relaxed_case(seed=9):
    tb = air5_tables(two_temperature=True)
    c = air5_box_case(10, 8, 6, seed=seed, tables=load_air5_tables(True), two_temperature=True)
    o = Oracle(c, tb)
    o.initialize()
    return c, o, tb


def test_vt_relaxation_conserves_energy_and_reaches_equilibrium():
    c, o, tb = relaxed_case()
    L = load()
    L.oc_ve_energy.restype = C.c_double
    L.oc_ve_energy.argtypes = [C.c_void_p, C.c_int, C.c_double]
    rhoE = interior(c, o.a["cv"])[4].copy()
    T0 = interior(c, o.a["bv"])[5].copy()
    tve0 = interior(c, o.a["tve"]).copy()
    y = interior(c, o.a["sv"])[:5].copy()
    o.op("vt_relaxation", dt=1.0)          # far longer than any relaxation time here
    assert np.array_equal(interior(c, o.a["cv"])[4], rhoE)
    eve = interior(c, o.a["sv"])[5]
    tp = C.cast(C.byref(tb), C.c_void_p)

    def e_el(i, T):           # the electronic part: E_ve minus the vibrational part
        th = tb.theta_v[i]
        vib = tb.gas_const[i] * th / math.expm1(th / T) if th > 0 and th / T < 200 else 0.0
        return L.oc_ve_energy(tp, i, T) - vib

    def e_vib(i, T):
        th = tb.theta_v[i]
        return tb.gas_const[i] * th / math.expm1(th / T) if th > 0 and th / T < 200 else 0.0

    for idx in [(1, 2, 3), (4, 0, 7), (5, 7, 9), (0, 3, 1)]:
        k, j, i = idx
        expect = sum(y[l][k, j, i] * (e_el(l, tve0[k, j, i]) + e_vib(l, T0[k, j, i])) for l in range(5))
        # the stored E_ve matches E_ve(Tve) to the Tve Newton tolerance (1e-8 relative)
        assert abs(eve[k, j, i] - expect) <= 1e-7 * abs(expect), idx
    # the translational temperature moved opposite to the vibrational energy gain
    T1 = interior(c, o.a["bv"])[5]
    hot = tve0 < T0
    assert np.all((T1 <= T0 + 1e-9)[hot])
    assert np.all((T1 >= T0 - 1e-9)[~hot])


@pytest.mark.parametrize("order", [2, 8])
def test_still_gas_viscous_energy_flux_is_the_eve_conduction(order):
    tb = air5_tables(two_temperature=True)
    c = air5_box_case(12, 10, 8, seed=10, tables=load_air5_tables(True), two_temperature=True,
                      ngg=4 if order == 8 else 3)
    c.run.update(viscous_order=order)
    c.bv[1:4] = 0.0
    c.bv[5] = 3000.0
    y0 = c.sv[:5, 0, 0, 0].copy()
    c.sv[:5] = y0[:, None, None, None]
    c.bv[0] = 1e4 / (c.bv[5] * (y0 * np.array(tb.gas_const[:5])).sum())
    c.bv[4] = 1e4
    o = Oracle(c, tb)
    o.initialize()
    o.op("zero_dq")
    o.op("viscous")
    dq = o.a["dq"]
    assert np.abs(dq[1:4]).max() <= 1e-12 * np.abs(dq[4]).max()
    assert np.abs(dq[5:10]).max() <= 1e-12 * np.abs(dq[4]).max()
    assert np.abs(dq[4]).max() > 0
    assert np.allclose(dq[4], dq[10], rtol=1e-12, atol=1e-12 * np.abs(dq[4]).max())


def test_reaction_temperature_exponents():
    t2 = air5_tables(two_temperature=True)
    for r in range(t2.n_reac):
        t2.tcf_a[r], t2.tcf_b[r], t2.tcb_a[r], t2.tcb_b[r] = 0.0, 1.0, 0.0, 1.0
    t1 = air5_tables(two_temperature=False)
    L = load()
    rho, T, Tve = 0.02, 7000.0, 4500.0
    y = np.ascontiguousarray([0.6, 0.1, 0.05, 0.15, 0.1] + [0.0] * 15)
    Ts = 0.5 * (Tve + 800.0 + math.sqrt((Tve - 800.0) ** 2 + 80.0 ** 2))
    out = []
    for tb, a, b in ((t2, T, Tve), (t1, Ts, Ts)):
        om, od = np.zeros(20), np.zeros(20)
        L.oc_chem_source(C.byref(tb), rho, a, b, y.ctypes.data_as(PD), om.ctypes.data_as(PD), od.ctypes.data_as(PD))
        out.append(om[:5].copy())
    assert np.allclose(out[0], out[1], rtol=1e-13, atol=1e-13 * np.abs(out[1]).max())
