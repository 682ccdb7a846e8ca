"""The oracle's characteristic WENO basis (schemes 52 / 72) pinned by the
algebra the eigen-decomposition must satisfy
(src/WENO.cu:1146-1457, 1872-2193, 2537-2857;
docs/two_temperature_weno_characteristic_derivation.md:549-778):

* for one state (left = right), R L = I: the back-projection inverts the
  projection, for the perfect gas, the 9-species H2/air mixture and the
  two-temperature AIR-5 mixture (the E_ve row), on both branches of the
  tangent basis (|k_x| > 0.8 and not);
* for two different states every entry of R L - I vanishes except the
  energy row's density column, which is the Roe state's enthalpy defect
  H~ - |u~|^2 / 2 - sum_l Y~_l h_tr,l(T~) - E_ve~ (for the perfect gas
  H~ - |u~|^2 / 2 - c_p T~): the reference's Roe average of H is not the
  enthalpy at its Roe temperature, and nothing else is off;
* on a Cartesian grid a uniform state's face flux equals the physical flux
  J k.F(Q) (the projection, the Lax-Friedrichs split and WENO of constant
  data give R L J k.F), and the divergence vanishes.

The matrices come from oc_char_eigen, which applies the projection and the
back-projection that char_weno_face uses to unit vectors."""
import ctypes as C

import numpy as np
import pytest

from corefl_amd.cases import air5_box_case, reacting_box_case, sod_case
from oracle import Oracle, air5_tables, h2_tables, load

R_AIR = 8314.462618 / 28.9647
GAMMA = 1.4


def _cases():
    c_air = sod_case(12, 7, scheme=52)
    rng = np.random.default_rng(5)
    c_air.bv[0] = rng.uniform(0.2, 2.0, c_air.bv[0].shape)
    c_air.bv[4] = rng.uniform(0.3, 3.0, c_air.bv[0].shape)
    c_air.bv[5] = c_air.bv[4] / (c_air.bv[0] * R_AIR)
    c_air.bv[1:4] = rng.uniform(-0.8, 0.8, c_air.bv[1:4].shape)
    c_h2 = reacting_box_case(8, 8, 8, seed=11, T_range=(300.0, 2600.0), u_rms=300.0)
    c_2t = air5_box_case(8, 8, 8, seed=12, two_temperature=True, T_range=(500.0, 12000.0), u_rms=2000.0)
    return [("air", c_air, None), ("h2", c_h2, h2_tables()), ("air5_2t", c_2t, air5_tables(True))]


@pytest.fixture(scope="module")
def states():
    out = []
    for name, case, tables in _cases():
        o = Oracle(case, tables)
        o.initialize()
        out.append((name, case, o))
    return out


def eigen(o, ql, pl, qr, pr, k):
    nv = o.run.n_var
    Lm, Rm = np.zeros(nv * nv), np.zeros(nv * nv)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ql, qr = np.ascontiguousarray(ql, float), np.ascontiguousarray(qr, float)
    load().oc_char_eigen(C.byref(o.run), C.byref(o.tables), P(ql), pl, P(qr), pr, *map(float, k), P(Lm), P(Rm))
    return Lm.reshape(nv, nv), Rm.reshape(nv, nv)


def node_state(o, idx):
    g = o.case.ngg
    k, j, i = g + idx[0], g + idx[1], g + idx[2]
    q = o.a["cv"][:, k, j, i].copy()
    return q, float(o.a["bv"][4, k, j, i])


DIRECTIONS = [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.9, 0.3, -0.2), (0.3, -0.7, 0.5),
              (-0.5, 0.2, 0.8)]


def test_one_state_right_inverts_left(states):
    rng = np.random.default_rng(1)
    for name, case, o in states:
        nv = o.run.n_var
        for _ in range(6):
            idx = tuple(int(v) for v in rng.integers(0, [case.mz, case.my, case.mx]))
            q, p = node_state(o, idx)
            for k in DIRECTIONS:
                Lm, Rm = eigen(o, q, p, q, p, k)
                scale = np.abs(Rm) @ np.abs(Lm)  # roundoff scale of each entry of R L
                err = np.abs(Rm @ Lm - np.eye(nv))
                assert np.all(err <= 1e-13 * scale + 1e-15), (name, k, err.max(), np.unravel_index(err.argmax(),
                                                                                                   err.shape))
                # and L R = I as well (the basis is square and complete)
                err2 = np.abs(Lm @ Rm - np.eye(nv))
                assert np.all(err2 <= 1e-13 * (np.abs(Lm) @ np.abs(Rm)) + 1e-15), (name, k, err2.max())


def _roe_defect(name, o, ql, pl, qr, pr):
    """H~ - |u~|^2/2 - h_tr(T~) (- E_ve~) at the Roe state of two nodes."""
    rl, rr = ql[0], qr[0]
    s = np.sqrt(rl * rr)
    wl, wr = 1.0 / (rl + s), 1.0 / (s + rr)
    u = wl * ql[1:4] + wr * qr[1:4]
    H = wl * (ql[4] + pl) + wr * (qr[4] + pr)
    pr_rho = wl * pl + wr * pr  # Roe average of p / rho
    if name == "air":
        T = pr_rho / R_AIR
        return H - 0.5 * u @ u - GAMMA * R_AIR * T / (GAMMA - 1)
    tb = o.tables
    ns = tb.n_spec
    y = wl * ql[5:] + wr * qr[5:]
    R = sum(y[l] * tb.gas_const[l] for l in range(ns))
    T = pr_rho / R
    h = np.zeros(ns)
    load().oc_enthalpy(C.byref(tb), T, h.ctypes.data_as(C.POINTER(C.c_double)))
    h_tr = h.copy()
    if name == "air5_2t":
        L = load()
        L.oc_ve_energy.restype = C.c_double
        L.oc_ve_energy.argtypes = [C.c_void_p, C.c_int, C.c_double]
        for l in range(ns):
            h_tr[l] -= L.oc_ve_energy(C.addressof(tb), l, T)
        return H - 0.5 * u @ u - y[:ns] @ h_tr - y[ns]
    return H - 0.5 * u @ u - y[:ns] @ h_tr


def test_two_states_differ_only_by_the_roe_enthalpy_defect(states):
    rng = np.random.default_rng(2)
    for name, case, o in states:
        nv = o.run.n_var
        worst_defect = 0.0
        for _ in range(6):
            ia = tuple(int(v) for v in rng.integers(0, [case.mz, case.my, case.mx]))
            ib = tuple(int(v) for v in rng.integers(0, [case.mz, case.my, case.mx]))
            ql, pl = node_state(o, ia)
            qr, pr = node_state(o, ib)
            for k in DIRECTIONS:
                Lm, Rm = eigen(o, ql, pl, qr, pr, k)
                M = Rm @ Lm - np.eye(nv)
                scale = np.abs(Rm) @ np.abs(Lm)
                defect = _roe_defect(name, o, ql, pl, qr, pr)
                # the energy row's density column carries the defect (times rho column of the identity)
                assert abs(M[4, 0] - defect) <= 1e-12 * scale[4, 0] + 1e-15, (name, M[4, 0], defect)
                M[4, 0] = 0.0
                assert np.all(np.abs(M) <= 1e-13 * scale + 1e-15), (name, np.abs(M).max())
                worst_defect = max(worst_defect, abs(defect) / scale[4, 0])
        # different states really were tested (the defect is not rounding)
        assert worst_defect > 1e-8, (name, worst_defect)


def _uniform(case, tables, seed):
    """The state of node (0, 0, 0) everywhere, with a seeded uniform velocity."""
    rng = np.random.default_rng(seed)
    g = case.ngg
    for a in range(6):
        case.bv[a] = case.bv[a][g, g, g]
    case.bv[1:4] = rng.uniform(-1, 1, 3)[:, None, None, None] * (0.5 if case.run["mixture"] == 0 else 300.0)
    for s in range(case.sv.shape[0]):
        case.sv[s] = case.sv[s][g, g, g]
    if case.tve is not None:
        # E_ve of the node's Tve, so that the periodic faces copy a set E_ve
        case.tve[:] = case.tve[g, g, g]
        L = load()
        L.oc_mixture_ve_energy.restype = C.c_double
        L.oc_mixture_ve_energy.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p]
        y = np.ascontiguousarray([case.sv[s][g, g, g] for s in range(tables.n_spec)])
        case.sv[tables.n_spec] = L.oc_mixture_ve_energy(C.addressof(tables), float(case.tve[g, g, g]),
                                                        y.ctypes.data, None)
    return case


def _physical_flux(o, case, dir_):
    """J k.F(Q) at every node from the oracle's initialised cv and p."""
    cv, p = o.a["cv"], o.a["bv"][4]
    m = o.a["metric"][3 * dir_:3 * dir_ + 3]
    U = (cv[1] * m[0] + cv[2] * m[1] + cv[3] * m[2]) / cv[0]
    F = np.empty_like(cv)
    F[0] = cv[0] * U
    for a in range(3):
        F[1 + a] = cv[1 + a] * U + p * m[a]
    F[4] = (cv[4] + p) * U
    F[5:] = cv[5:] * U
    return F * o.a["jac"]


@pytest.mark.parametrize("scheme", [52, 72])
@pytest.mark.parametrize("physics", ["air", "h2", "air5_2t"])
def test_uniform_state_face_flux_is_the_physical_flux(scheme, physics):
    if physics == "air":
        case, tables = sod_case(14, 7, scheme=scheme), None
    elif physics == "h2":
        tables = h2_tables()
        case = reacting_box_case(10, 9, 9, seed=3, amp=0.0)
    else:
        tables = air5_tables(True)
        case = air5_box_case(10, 9, 9, seed=4, two_temperature=True, amp=0.0, ngg=4)
    case.run["inviscid_scheme"] = scheme
    _uniform(case, tables, 9)
    o = Oracle(case, tables)
    o.initialize()
    dirs = (0, 1) if case.mz == 1 else (0, 1, 2)
    g = case.ngg
    for d, name, fl in zip(dirs, ("x", "y", "z"), ("ff", "gf", "hf")):
        o.op("zero_dq")
        o.op("inviscid_" + name)
        F = _physical_flux(o, case, d)
        # faces i+1/2 for i = -1 .. n-1 along the sweep: the flux plane at the left node
        sl = [slice(g, g + case.mz), slice(g, g + case.my), slice(g, g + case.mx)]
        sl[2 - d] = slice(g - 1, g + [case.mx, case.my, case.mz][d])
        got = o.a[fl][(slice(None),) + tuple(sl)]
        ref = F[(slice(None),) + tuple(sl)]
        scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
        scale = np.maximum(scale, np.abs(F[:5]).max())  # rows that vanish (zero velocity component) in flux units
        err = np.abs(got - ref).reshape(ref.shape[0], -1).max(axis=1) / scale
        assert np.all(err <= 1e-12), (physics, scheme, name, err)
        assert np.abs(o.a["dq"]).max() <= 1e-12 * np.abs(F).max() / np.abs(o.a["jac"]).max() * 1e3
