"""The oracle's two-temperature energy modes and V-T relaxation time
(src/Thermo.cu:635-901) pinned by closed forms and published correlations:

* harmonic-oscillator vibrational energy: its high-temperature expansion
  R (T - theta/2 + theta^2 / (12 T)) and its vanishing at low temperature;
* the electronic energy of a two-level species against the Boltzmann closed form;
* c_v,ve = dE_ve/dT for every AIR-5 species;
* E_tr(T) + E_ve(T) = sum Y (h - R T): the split of the equilibrium energy;
* the Tve Newton inversion returns the temperature an energy was made from;
* the Millikan-White time of pure N2 at 1 atm against the published
  correlation p tau = exp(1.16e-3 mu^1/2 theta^4/3 (T^-1/3 - 0.015 mu^1/4) - 18.42);
* tau proportional to 1/p at fixed temperature (Millikan-White and Park parts)."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import PD, air5_tables, load

D = C.c_double


@pytest.fixture(scope="module")
def tb():
    return air5_tables(two_temperature=True)


@pytest.fixture(scope="module")
def lib():
    L = load()
    tp = C.c_void_p
    L.oc_ve_energy.restype = D
    L.oc_ve_energy.argtypes = [tp, C.c_int, D]
    L.oc_ve_cv.restype = D
    L.oc_ve_cv.argtypes = [tp, C.c_int, D]
    L.oc_mixture_ve_energy.restype = D
    L.oc_mixture_ve_energy.argtypes = [tp, D, PD, PD]
    L.oc_mixture_tr_energy.restype = D
    L.oc_mixture_tr_energy.argtypes = [tp, D, PD, PD, PD, PD]
    L.oc_invert_tve.restype = D
    L.oc_invert_tve.argtypes = [tp, D, PD, D]
    L.oc_vt_source.restype = D
    L.oc_vt_source.argtypes = [tp, D, D, D, PD, PD, PD]
    return L


def P(t):
    return C.cast(C.byref(t), C.c_void_p)


def arr(x):
    x = np.ascontiguousarray(x, dtype=float)
    return x, x.ctypes.data_as(PD)


def test_vibrational_energy_limits(tb, lib):
    # N2 (theta_v 3395 K) with its electronic levels removed
    t = air5_tables(two_temperature=True)
    t.n_elec[0] = 0
    R, th = t.gas_const[0], t.theta_v[0]
    T = 2.0e5
    e = lib.oc_ve_energy(P(t), 0, T)
    assert abs(e - R * (T - th / 2 + th * th / (12 * T))) <= 1e-9 * R * T
    assert lib.oc_ve_energy(P(t), 0, 100.0) <= 1e-12 * R * 100.0
    # atoms have no vibrational mode
    t.n_elec[3] = 0
    assert lib.oc_ve_energy(P(t), 3, 5000.0) == 0.0


def test_two_level_electronic_energy(tb, lib):
    t = air5_tables(two_temperature=True)
    i = 4                                   # O: replace its levels by a two-level system
    t.theta_v[i] = 0.0
    t.n_elec[i] = 2
    t.elec_theta[i][0], t.elec_g[i][0] = 0.0, 1.0
    t.elec_theta[i][1], t.elec_g[i][1] = 1000.0, 3.0
    R = t.gas_const[i]
    for T in (150.0, 700.0, 4000.0):
        x = 3.0 * math.exp(-1000.0 / T)
        assert abs(lib.oc_ve_energy(P(t), i, T) - R * 1000.0 * x / (1.0 + x)) <= 1e-12 * R * 1000.0
        # c_v of a two-level system: R (theta/T)^2 x / (1 + x)^2
        assert abs(lib.oc_ve_cv(P(t), i, T) - R * (1000.0 / T) ** 2 * x / (1 + x) ** 2) <= 1e-10 * R


@pytest.mark.parametrize("T", [300.0, 1500.0, 6000.0, 15000.0])
def test_cv_is_the_derivative_of_the_energy(tb, lib, T):
    d = 1e-4 * T
    for i in range(5):
        fd = (lib.oc_ve_energy(P(tb), i, T + d) - lib.oc_ve_energy(P(tb), i, T - d)) / (2 * d)
        cv = lib.oc_ve_cv(P(tb), i, T)
        assert abs(fd - cv) <= 1e-6 * max(abs(cv), tb.gas_const[i] * 1e-3), (i, fd, cv)


def test_equilibrium_energy_split(tb, lib):
    y, py = arr([0.6, 0.15, 0.05, 0.12, 0.08] + [0.0] * 15)
    for T in (400.0, 3000.0, 12000.0):
        e_tr = lib.oc_mixture_tr_energy(P(tb), T, py, None, None, None)
        e_ve = lib.oc_mixture_ve_energy(P(tb), T, py, None)
        h, ph = arr(np.zeros(20))
        load().oc_enthalpy(C.byref(tb), T, ph)
        e_eq = sum(y[i] * (h[i] - tb.gas_const[i] * T) for i in range(5))
        assert abs(e_tr + e_ve - e_eq) <= 1e-12 * abs(e_eq) + 1e-6


def test_tve_inversion_round_trip(tb, lib):
    rng = np.random.default_rng(4)
    for _ in range(20):
        y = rng.uniform(0.01, 1.0, 5)
        y /= y.sum()
        y, py = arr(list(y) + [0.0] * 15)
        tve = rng.uniform(250.0, 20000.0)
        e = lib.oc_mixture_ve_energy(P(tb), tve, py, None)
        back = lib.oc_invert_tve(P(tb), e, py, rng.uniform(300.0, 20000.0))
        assert abs(back - tve) <= 1e-5 * tve, (tve, back)


def test_millikan_white_time_of_nitrogen(tb, lib):
    y, py = arr([1.0] + [0.0] * 19)
    R = tb.gas_const[0]
    mu = 28.0134 / 2
    for T in (1500.0, 2500.0, 4000.0):
        rho = 101325.0 / (R * T)                  # 1 atm
        eq, peq = arr([0.0])
        tau, ptau = arr([0.0])
        lib.oc_vt_source(P(tb), rho, T, 0.7 * T, py, peq, ptau)
        a = 1.16e-3 * math.sqrt(mu) * 3395.0 ** (4.0 / 3.0)
        b = 0.015 * mu ** 0.25
        tau_mw = math.exp(a * (T ** (-1.0 / 3.0) - b) - 18.42)
        # the file's pair coefficients (221, 0.029) restate the correlation to a few percent;
        # Park's correction adds well under 1% at these temperatures
        assert abs(tau[0] / tau_mw - 1.0) < 0.12, (T, tau[0], tau_mw)


def test_relaxation_time_scales_with_inverse_pressure(tb, lib):
    y, py = arr([0.7, 0.2, 0.02, 0.05, 0.03] + [0.0] * 15)
    for T in (2000.0, 9000.0):
        out = []
        for rho in (0.01, 0.02):
            eq, peq = arr([0.0])
            tau, ptau = arr([0.0])
            src = lib.oc_vt_source(P(tb), rho, T, 0.6 * T, py, peq, ptau)
            out.append((src, tau[0]))
        assert abs(out[1][1] / out[0][1] - 0.5) < 1e-12
        assert abs(out[1][0] / out[0][0] - 4.0) < 1e-11
