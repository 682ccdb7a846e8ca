"""The oracle's primitive-variable inviscid path (reconstruction + LF / AUSM+ /
HLLC, and Roe; src/InviscidScheme.cu:28-219, 221-395, 397-787,
src/Reconstruction.cu) pinned by consistency (a uniform state gives the
physical flux J k.F(Q) at every face), upwinding (a supersonic flow takes its
face flux from the upwind state alone), exactness of MUSCL/NND2 and of the
WENO5 / WENO7 reconstructions (4 / 5, no reference counterpart) on linear
data, and the Sod shock tube against the reference's exact solution
(example/sodShockTube/exact.dat)."""
from pathlib import Path

import numpy as np
import pytest

from corefl_amd.cases import multiblock_case, sod_case
from oracle import Oracle

EXACT = Path(__file__).parent / "golden" / "sod" / "exact.dat"
SOLVERS = {1: "LF", 2: "Roe", 3: "AUSM+", 4: "HLLC"}


def cartesian(physics, scheme, reconstruction, n=(12, 10, 9)):
    c = multiblock_case(split=(1, 1, 1), widths=[[n[0]], [n[1]], [n[2]]], seed=2, physics=physics, rotate=False,
                        wiggle=0.0, ngg=4 if reconstruction == 5 or physics == "h2" else 3)[0]
    c.run.update(inviscid_scheme=scheme, reconstruction=reconstruction, limiter=0, viscous_order=0, reaction=0)
    return c


def physical_flux(case, o, q, sv, i, j, k):
    """J (k.F) at node (i,j,k) for primitive state q = (rho,u,v,w,p,T), x sweep."""
    g = case.ngg
    m = o.a["metric"][0:3, k + g, j + g, i + g]
    J = o.a["jac"][k + g, j + g, i + g]
    rho, u, v, w, p, T = q
    U = m[0] * u + m[1] * v + m[2] * w
    if case.run["mixture"]:
        h = np.zeros(9)
        import ctypes as C
        from oracle import PD
        o.lib.oc_enthalpy(C.byref(o.tables), max(T, 1.0), h.ctypes.data_as(PD))
        # e_tr at the state's own temperature T = p / (rho R)
        R = (sv * np.array(o.tables.gas_const[:9])).sum()
        Tn = p / (rho * R)
        o.lib.oc_enthalpy(C.byref(o.tables), max(Tn, 1.0), h.ctypes.data_as(PD))
        E = rho * (0.5 * (u * u + v * v + w * w) + (sv * (h - np.array(o.tables.gas_const[:9]) * max(Tn, 1.0))).sum())
    else:
        E = rho * 0.5 * (u * u + v * v + w * w) + p / 0.4
    f = [rho * U, rho * u * U + p * m[0], rho * v * U + p * m[1], rho * w * U + p * m[2], (E + p) * U]
    f += list(rho * U * sv)
    return J * np.array(f)


@pytest.mark.parametrize("physics,scheme", [("air", 1), ("air", 2), ("air", 3), ("air", 4), ("h2", 2), ("h2", 3),
                                            ("h2", 4)])
@pytest.mark.parametrize("rec", [1, 2, 4, 5])
def test_uniform_state_gives_the_physical_flux(physics, scheme, rec, h2_oracle_tables):
    case = cartesian(physics, scheme, rec)
    q = case.bv[:, case.ngg + 2, case.ngg + 2, case.ngg + 2].copy()
    q[1:4] = [170.0, -40.0, 25.0] if physics == "h2" else [0.6, -0.2, 0.1]
    case.bv[:] = q[:, None, None, None]
    sv = case.sv[:, 0, 0, 0].copy()
    if case.run["n_scalar"]:
        case.sv[:] = sv[:, None, None, None]
    o = Oracle(case, h2_oracle_tables if physics == "h2" else None)
    o.op("properties")
    o.op("zero_dq")
    o.op("inviscid_x")
    g = case.ngg
    for (i, j, k) in [(3, 4, 5), (0, 2, 1), (case.mx - 2, case.my - 1, 3)]:
        fc = o.a["ff"][:, k + g, j + g, i + g]
        ref = physical_flux(case, o, q, sv, i, j, k)
        scale = np.abs(ref).max()
        assert np.allclose(fc, ref, rtol=0, atol=1e-12 * scale), (SOLVERS[scheme], fc, ref)
    assert np.abs(o.a["dq"]).max() <= 1e-12 * np.abs(o.a["ff"]).max()


@pytest.mark.parametrize("scheme", [3, 4])
def test_supersonic_flow_is_upwinded(scheme):
    case = cartesian("air", scheme, 1)
    g = case.ngg
    x = np.arange(-g, case.mx + g)[None, None, :] * np.ones(case.padded_shape)
    case.bv[0] = 1.0 + 0.1 * np.sin(0.7 * x)
    case.bv[1] = 3.0 + 0.05 * np.cos(0.5 * x)     # Mach > 2 everywhere
    case.bv[2] = 0.1
    case.bv[3] = -0.05
    case.bv[4] = 1.0 / 1.4 + 0.05 * np.sin(0.3 * x)
    o = Oracle(case, None)
    o.op("zero_dq")
    o.op("inviscid_x")
    for i in range(0, case.mx - 1, 3):
        q = case.bv[:, g + 2, g + 3, g + i]
        fc = o.a["ff"][:, g + 2, g + 3, g + i]
        ref = physical_flux(case, o, q, np.zeros(0), i, 3, 2)
        assert np.allclose(fc, ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max()), SOLVERS[scheme]


@pytest.mark.parametrize("rec", [2, 3, 4, 5])
def test_muscl_nnd_and_weno_are_exact_on_linear_data(rec):
    """A linear density profile carried supersonically: the face mass flux is
    rho at the face midpoint times U J (slopes of linear data are equal, so
    the limiter keeps them; every WENO candidate is exact for linear data and
    the smoothness indicators are equal, so the weights are the linear ones)."""
    case = cartesian("air", 3, rec)
    g = case.ngg
    x = np.arange(-g, case.mx + g)[None, None, :] * np.ones(case.padded_shape)
    case.bv[0] = 1.0 + 0.01 * x
    case.bv[1], case.bv[2], case.bv[3] = 3.0, 0.0, 0.0
    case.bv[4] = 1.0 / 1.4
    o = Oracle(case, None)
    o.op("zero_dq")
    o.op("inviscid_x")
    m = o.a["metric"][0, g + 1, g + 1, g + 1]
    J = o.a["jac"][g + 1, g + 1, g + 1]
    for i in range(-1, case.mx - 1):
        rho_face = 1.0 + 0.01 * (i + 0.5)
        assert abs(o.a["ff"][0, g + 1, g + 1, g + i] - rho_face * 3.0 * m * J) <= 1e-13 * rho_face * 3.0 * m * J


# L1 errors against the exact solution at t = 0.2 (400 nodes, CFL 0.5):
# first order about 8e-3 in rho, MUSCL about 2.2e-3, WENO5 / WENO7 reconstruction
# about 1.1e-3; bounds at 1.2x
SOD_BOUNDS = {(1, 1): (1.45e-2, 2.1e-2), (2, 1): (1.0e-2, 1.55e-2), (3, 1): (9.4e-3, 1.4e-2), (4, 1): (1.0e-2, 1.6e-2),
              (1, 2): (3.6e-3, 5.1e-3), (2, 2): (2.8e-3, 4.5e-3), (3, 2): (2.6e-3, 4.2e-3), (4, 2): (2.8e-3, 4.7e-3),
              (3, 4): (1.3e-3, 2.5e-3), (4, 4): (1.55e-3, 2.9e-3), (3, 5): (1.35e-3, 2.35e-3),
              (4, 5): (1.25e-3, 2.25e-3)}


@pytest.mark.parametrize("scheme,rec", sorted(SOD_BOUNDS))
def test_sod_primitive_path_matches_exact_solution(scheme, rec):
    case = sod_case(400, 7, 72 if rec == 5 else 52, 0.5)
    case.run.update(inviscid_scheme=scheme, reconstruction=rec, limiter=0)
    o = Oracle(case, None)
    o.initialize()
    t = 0.0
    while t < 0.2:
        t += o.step(0.0, fixed=False)
    g = case.ngg
    x = case.x[0, g, g + 3, g:g + case.mx]
    ex = np.array(EXACT.read_text().split()[1:], dtype=float).reshape(-1, 4)
    rho = o.a["bv"][0, g, g + 3, g:g + case.mx]
    u = o.a["bv"][1, g, g + 3, g:g + case.mx]
    b_rho, b_u = SOD_BOUNDS[(scheme, rec)]
    assert np.abs(rho - np.interp(x, ex[:, 0], ex[:, 2])).mean() < b_rho, SOLVERS[scheme]
    assert np.abs(u - np.interp(x, ex[:, 0], ex[:, 1])).mean() < b_u, SOLVERS[scheme]
