"""The oracle's component-wise WENO (inviscid_scheme 51 / 71), its hybrid
linear/WENO switch, the species positivity limiter and the shock sensors
(src/WENO.cu:831-1086, 1459-1527, 3033-3196; src/FieldOperation.cu:122-228),
pinned by what the mathematics fixes:

* on a smooth 1-D Euler field the flux divergence dq / J converges to
  -d(f)/dx at 5th (UD5) / 7th (UD7) order with the linear weights, and the
  WENO weights keep that order on 40-80 nodes per wave; a wrong stencil
  coefficient or an index shift between the split halves breaks the order;
* mirroring the state (x -> -x, u -> -u) mirrors dq with the x-momentum row
  negated, for the component-wise and the characteristic schemes;
* a sensor above the threshold everywhere gives the plain WENO result bit for
  bit;
* the positivity limiter leaves every face whose two one-sided updates stay
  non-negative untouched, never makes a violated side worse, and brings the
  violated side exactly to zero wherever it changes a flux (theta in (0, 1));
* Ducros vanishes for a rotation and for an expansion and tends to one for a
  pure compression; Jameson vanishes for linear pressure and has the closed
  form |2p_1 - 2p_0| / (p_0 + p_1 + 2p) at a jump; the jump sensor is the
  largest relative neighbour jump of rho and p."""
import numpy as np
import pytest

from corefl_amd import reacting_box_case, sod_case
from oracle import Oracle, h2_tables

GAMMA = 1.4


def smooth_air(nx, scheme, u=0.5):
    """sod_case grid with a smooth density wave at constant u and p on the padded extent."""
    c = sod_case(nx, 7, scheme=scheme)
    x = c.x[0]
    c.bv[0] = 1.0 + 0.2 * np.sin(2 * np.pi * x)
    c.bv[1] = u
    c.bv[2] = c.bv[3] = 0.0
    c.bv[4] = 1.0
    c.bv[5] = c.bv[4] * 28.9647 / (c.bv[0] * 8314.462618)
    return c


def dq_after(o, case, dirs=("x",), reset=True):
    if reset:
        o.a["bv"][:] = case.bv
        o.op("cv_from_bv")
    o.op("zero_dq")
    for d in dirs:
        o.op("inviscid_" + d)
    return o.a["dq"].copy()


def exact_divergence(case, u=0.5):
    """-d/dx of the Euler flux of the smooth wave (rho' u, rho' u^2, 0, 0, rho' u^3 / 2)."""
    g = case.ngg
    x = case.x[0][g:g + case.mz, g:g + case.my, g:g + case.mx]
    drho = 0.2 * 2 * np.pi * np.cos(2 * np.pi * x)
    return -np.stack([drho * u, drho * u * u, 0 * x, 0 * x, 0.5 * drho * u ** 3])


@pytest.mark.parametrize("scheme,order", [(51, 5), (71, 7)])
@pytest.mark.parametrize("linear", [True, False])
def test_component_wise_convergence(scheme, order, linear):
    errs = []
    for nx in (41, 81):
        c = smooth_air(nx, scheme)
        if linear:
            c.run.update(sensor_threshold=0.5)  # sensor plane stays 0: every face takes the linear weights
        o = Oracle(c, None)
        o.initialize()
        dq = dq_after(o, c)
        g = c.ngg
        jac = c.jac[g:g + c.mz, g:g + c.my, g:g + c.mx]
        ex = exact_divergence(c)
        sl = (slice(None), slice(None), slice(None), slice(8, nx - 8))  # away from the outflow copies
        errs.append(np.abs(dq / jac - ex)[sl][[0, 1, 4]].max())
    rate = np.log2(errs[0] / errs[1])
    assert rate > order - 0.4, (errs, rate)


def mirrored(case):
    m = sod_case(case.mx, case.my, scheme=case.run["inviscid_scheme"])
    m.bv = case.bv[..., ::-1].copy()
    m.bv[1] *= -1
    m.sv = case.sv[..., ::-1].copy()
    return m


@pytest.mark.parametrize("scheme", [51, 71, 52, 72])
def test_mirror_symmetry(scheme):
    rng = np.random.default_rng(3)
    c = sod_case(40, 7, scheme=scheme)
    c.bv[0] = rng.uniform(0.5, 2.0, c.bv[0].shape[-1])[None, None, :] * np.ones_like(c.bv[0])
    c.bv[1] = rng.uniform(-0.5, 0.5, c.bv[0].shape[-1])[None, None, :] * np.ones_like(c.bv[0])
    c.bv[4] = rng.uniform(0.5, 2.0, c.bv[0].shape[-1])[None, None, :] * np.ones_like(c.bv[0])
    c.bv[5] = c.bv[4] * 28.9647 / (c.bv[0] * 8314.462618)
    m = mirrored(c)
    oc, om = Oracle(c, None), Oracle(m, None)
    oc.initialize()
    om.initialize()
    a = dq_after(oc, c)
    b = dq_after(om, m)[..., ::-1]
    b[1] *= -1
    scale = np.abs(a).max(axis=(1, 2, 3))
    err = np.abs(a - b).max(axis=(1, 2, 3))
    assert np.all(err[[0, 1, 4]] <= 1e-12 * scale[[0, 1, 4]]), err / np.maximum(scale, 1e-300)


@pytest.mark.parametrize("scheme", [71, 72])
def test_sensor_above_threshold_is_plain_weno(scheme):
    c = reacting_box_case(12, 10, 8, seed=5)
    c.run.update(inviscid_scheme=scheme)
    tb = h2_tables()
    o1 = Oracle(c, tb)
    o1.initialize()
    ref = dq_after(o1, c, dirs=("x", "y", "z"), reset=False)
    c.run.update(sensor_threshold=0.3)
    o2 = Oracle(c, tb)
    o2.initialize()
    o2.a["shock_sensor"][:] = 1.0
    hyb = dq_after(o2, c, dirs=("x", "y", "z"), reset=False)
    assert np.array_equal(ref, hyb)
    o2.a["shock_sensor"][:] = 0.0
    lin = dq_after(o2, c, dirs=("x", "y", "z"), reset=False)
    assert not np.array_equal(ref, lin)


def species_front_case(seed=11, scheme=71):
    """H2/air box with the mixture fraction thresholded to 0 / 1: H2 vanishes
    where O2 and N2 are present and the other way round, so high-order
    species fluxes overshoot next to the fronts."""
    c = reacting_box_case(16, 12, 10, seed=seed)
    z = (c.sv[0] > 0.5).astype(float)
    c.sv[:] = 0.0
    c.sv[0] = z
    c.sv[2] = 0.233 * (1 - z)
    c.sv[8] = 0.767 * (1 - z)
    mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
    imw = (c.sv / mw[:, None, None, None]).sum(axis=0)
    c.bv[0] = c.bv[4] / (8314.462618 * imw * c.bv[5])
    c.run.update(inviscid_scheme=scheme, fixed_time_step=1)
    return c


def one_sided(o, c, dt, alpha=1.0 / 3.0):
    """alpha/2 J rhoY_left - dt F and alpha/2 J rhoY_right + dt F at every x face (i = -1 .. mx-2 here)."""
    g = c.ngg
    cv = o.a["cv"][5:]
    J = o.a["jac"]
    F = o.a["ff"][5:]
    sl = (slice(None), slice(g, g + c.mz), slice(g, g + c.my))
    L = slice(g - 1, g + c.mx - 1)
    R = slice(g, g + c.mx)
    up = 0.5 * alpha * J[sl[1:] + (L,)] * cv[sl + (L,)] - dt * F[sl + (L,)]
    um = 0.5 * alpha * J[sl[1:] + (R,)] * cv[sl + (R,)] + dt * F[sl + (L,)]
    return up, um, F[sl + (L,)].copy()


@pytest.mark.parametrize("scheme", [71, 72])
def test_positivity_limiter_properties(scheme):
    tb = h2_tables()
    c = species_front_case(scheme=scheme)
    o0 = Oracle(c, tb)
    o0.initialize()
    dt = 1.5 * o0.cfl_dt()   # a step large enough that some one-sided updates turn negative
    c.run.update(dt=dt)
    o0 = Oracle(c, tb)
    o0.initialize()
    o0.op("zero_dq")
    o0.op("inviscid_x")
    up0, um0, f0 = one_sided(o0, c, dt)
    c.run.update(positive_preserving=1)
    o1 = Oracle(c, tb)
    o1.initialize()
    o1.op("zero_dq")
    o1.op("inviscid_x")
    up1, um1, f1 = one_sided(o1, c, dt)
    ok = (up0 >= 0) & (um0 >= 0)
    assert np.array_equal(f0[ok], f1[ok])            # untouched where nothing is violated
    changed = f0 != f1
    assert changed.sum() > 10                         # the limiter acted
    assert np.all(changed <= ~ok)
    scale = 0.5 / 3 * np.abs(o0.a["cv"][5:]).max() * np.abs(o0.a["jac"]).max()
    tol = 1e-9 * scale
    viol_p, viol_m = up0 < 0, um0 < 0
    assert np.all(up1[viol_p & changed] >= up0[viol_p & changed] - tol)
    assert np.all(um1[viol_m & changed] >= um0[viol_m & changed] - tol)
    zeroed = (np.abs(up1) <= tol) & viol_p | (np.abs(um1) <= tol) & viol_m
    assert np.array_equal(zeroed & changed, changed)


def cartesian_sensor_case(u, v, w, p, rho=None):
    c = sod_case(12, 10, scheme=52)
    x, y = c.x[0], c.x[1]
    c.bv[1], c.bv[2], c.bv[3] = u(x, y), v(x, y), w(x, y)
    c.bv[4] = p(x, y)
    c.bv[0] = rho(x, y) if rho is not None else np.ones_like(x)
    c.bv[5] = c.bv[4] * 28.9647 / (c.bv[0] * 8314.462618)
    return c


def sensor(c, kind, eps=1e-30):
    c.run.update(shock_sensor=kind, sensor_eps=eps, sensor_threshold=0.1, inviscid_scheme=52)
    o = Oracle(c, None)
    o.a["bv"][:] = c.bv
    o.op("shock_sensor")
    g = c.ngg
    return o.a["shock_sensor"][g:g + c.mz, g:g + c.my, g:g + c.mx]


def test_ducros_sensor_limits():
    zero = lambda x, y: 0 * x
    one = lambda x, y: 1 + 0 * x
    rot = sensor(cartesian_sensor_case(lambda x, y: -3 * y, lambda x, y: 3 * x, zero, one), 0)
    assert np.abs(rot).max() <= 1e-20
    expa = sensor(cartesian_sensor_case(lambda x, y: 2 * x, lambda x, y: 2 * y, zero, one), 0)
    assert np.all(expa == 0)
    comp = sensor(cartesian_sensor_case(lambda x, y: -2 * x, lambda x, y: -5 * y, zero, one), 0)
    assert np.all(np.abs(comp - 1) <= 1e-12)
    # a compression with an equal rotation: div^2 / (div^2 + curl^2) = 1/2
    half = sensor(cartesian_sensor_case(lambda x, y: -2 * x - y, lambda x, y: x, zero, one), 0)
    assert np.all(np.abs(half - 0.5) <= 1e-12)


def test_jameson_and_jump_sensors():
    zero = lambda x, y: 0 * x
    lin = sensor(cartesian_sensor_case(zero, zero, zero, lambda x, y: 1 + 0.3 * x + 0.2 * y), 1)
    assert np.abs(lin).max() <= 1e-14
    c = cartesian_sensor_case(zero, zero, zero, lambda x, y: np.where(x > 0.5, 2.0, 1.0),
                              rho=lambda x, y: np.where(x > 0.5, 1.5, 1.0))
    s1 = sensor(c, 1)
    s2 = sensor(c, 2)
    g = c.ngg
    x = c.x[0][g:g + c.mz, g:g + c.my, g:g + c.mx]
    dx = x[0, 0, 1] - x[0, 0, 0]
    left = (x <= 0.5) & (x + dx > 0.5)     # last node before the jump
    right = (x > 0.5) & (x - dx <= 0.5)    # first node after it
    # Jameson: |p_{i+1} + p_{i-1} - 2p| / (p_{i+1} + p_{i-1} + 2p), j and k terms zero
    assert np.allclose(s1[left], 1.0 / 5.0, rtol=1e-14)
    assert np.allclose(s1[right], 1.0 / 7.0, rtol=1e-14)
    assert np.all(s1[~(left | right)] == 0)
    # jump sensor looks forward only: max(|rho_{i+1} - rho| / rho, |p_{i+1} - p| / p)
    assert np.allclose(s2[left], 1.0, rtol=1e-14)
    assert np.all(s2[~left] == 0)
