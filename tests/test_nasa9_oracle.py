"""The oracle's NASA-9 thermodynamics (src/Thermo.cu nasa_7_or_9 == 9 branches)
pinned by published values and by consistency:

* standard enthalpies of formation and entropies at 298.15 K of N2, O2, NO, N
  and O (the NASA Glenn / Gurvich values the fits are built from);
* cp of N2, O2, N (exactly 5R/2) and O (with its fine-structure levels) at 298.15 K;
* cp = dh/dT inside every range, h and cp continuous at the range cuts;
* G/RT = H/RT - S/R, with S from the fit's b2 term;
* linear enthalpy extrapolation outside [200 K, 20000 K] with the bound's cp."""
import ctypes as C

import numpy as np
import pytest

from oracle import PD, air5_tables, load

R_U = 8314.462618
SPECIES = ["N2", "O2", "NO", "N", "O"]
# H_f(298.15) in kJ/mol and S(298.15) in J/(mol K)
HF = {"N2": 0.0, "O2": 0.0, "NO": 91.271, "N": 472.680, "O": 249.175}
S298 = {"N2": 191.609, "O2": 205.148, "NO": 210.758, "N": 153.301, "O": 161.059}


@pytest.fixture(scope="module")
def tb():
    return air5_tables(two_temperature=False)


def h_cp(tb, t):
    lib = load()
    h = np.zeros(20)
    cp = np.zeros(20)
    lib.oc_enthalpy_and_cp(C.byref(tb), t, h.ctypes.data_as(PD), cp.ctypes.data_as(PD))
    return h[:5].copy(), cp[:5].copy()


def gibbs(tb, t):
    g = np.zeros(20)
    load().oc_gibbs_div_rt(C.byref(tb), t, g.ctypes.data_as(PD))
    return g[:5].copy()


def test_formation_enthalpies_and_cp(tb):
    h, cp = h_cp(tb, 298.15)
    mw = np.array(tb.mw[:5])
    h_molar = h * mw / 1e6          # J/kg * kg/kmol -> kJ/mol
    cp_molar = cp * mw / 1e3        # J/(mol K)
    for i, s in enumerate(SPECIES):
        assert abs(h_molar[i] - HF[s]) < 0.01, s
    assert abs(cp_molar[0] - 29.124) < 0.01
    assert abs(cp_molar[1] - 29.376) < 0.01
    assert abs(cp_molar[3] - 20.786) < 1e-3   # N: 5R/2, no electronic level within reach at 298 K
    assert abs(cp_molar[4] - 21.912) < 1e-2   # O: 5R/2 plus its 227 K / 326 K fine-structure levels


def test_gibbs_is_h_minus_ts(tb):
    t = 298.15
    h, _ = h_cp(tb, t)
    g = gibbs(tb, t)
    gas = np.array(tb.gas_const[:5])
    mw = np.array(tb.mw[:5])
    for i, s in enumerate(SPECIES):
        s_over_r = S298[s] / (R_U / 1e3)
        assert abs(g[i] - (h[i] / (gas[i] * t) - s_over_r)) < 2e-3, s
    # and at a high temperature, the same identity with S from the fit itself:
    # d(G/RT)/dT = -H/(R T^2)
    for t in (800.0, 3000.0, 9000.0, 15000.0):
        d = 1e-3 * t
        dg = (gibbs(tb, t + d) - gibbs(tb, t - d)) / (2 * d)
        h, _ = h_cp(tb, t)
        assert np.allclose(dg, -h / (gas * t * t), rtol=1e-6, atol=1e-12), t


@pytest.mark.parametrize("t", [250.0, 700.0, 2500.0, 5000.0, 8000.0, 18000.0])
def test_cp_is_dh_dt(tb, t):
    d = 1e-4 * t
    hp, _ = h_cp(tb, t + d)
    hm, _ = h_cp(tb, t - d)
    _, cp = h_cp(tb, t)
    assert np.allclose((hp - hm) / (2 * d), cp, rtol=1e-7)


@pytest.mark.parametrize("cut", [1000.0, 6000.0])
def test_continuity_at_cuts(tb, cut):
    hl, cl = h_cp(tb, cut * (1 - 1e-12))
    hr, cr = h_cp(tb, cut * (1 + 1e-12))
    scale = np.abs(np.array(tb.gas_const[:5])) * cut
    assert np.all(np.abs(hl - hr) < 2e-4 * scale)
    assert np.allclose(cl, cr, rtol=2e-3)


def test_extrapolation_outside_the_fit(tb):
    for lo, t in ((200.0, 150.0), (20000.0, 25000.0)):
        hb, cb = h_cp(tb, lo)
        h, cp = h_cp(tb, t)
        assert np.allclose(cp, cb, rtol=0, atol=0)
        assert np.allclose(h, hb + cb * (t - lo), rtol=1e-14)
        assert np.array_equal(gibbs(tb, t), gibbs(tb, lo))
