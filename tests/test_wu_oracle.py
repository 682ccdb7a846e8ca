"""The oracle's Wu operator splitting (temporal_scheme 4,
src/StrangSplitting.cu:15-865) pinned by what the mathematics fixes:

* on a uniform state the transport part vanishes (Cn = 0 and dq - Cn = 0 in
  the RK sub-step), so a step is the reaction sub-step alone; with the linear
  chain A -> B -> C (constant rates, species of identical thermodynamics) each
  chemistry sub-step of length h is a closed-form matrix: I + hK (explicit,
  chem_advance_method 0), I + h (I - hK)^-1 K (point-implicit with the exact
  Jacobian, 1) and the SSP-RK3 polynomial I + hK + (hK)^2/2 + (hK)^3/6 (3);
  a step splits dt by the point's chemical time scale ts into
  n_sub = floor(dt / ts) sub-steps of dt / (n_sub + 1) and a last one of the
  rest: the first step by a preset ts = dt / 2.5, the second by the time
  scale 1/k1 the first one left (both three sub-steps of dt / 3);
* on a smooth reacting field the split step converges to the unsplit SSP-RK3
  solution as dt shrinks (first order: the error halves with dt)."""
import numpy as np
import pytest

from corefl_amd.cases import reacting_box_case
from oracle import Oracle, h2_tables

K1, K2 = 2.5e7, 4.0e6  # 1/s


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


def uniform_chain_case(y0, T=1200.0, p=2e5):
    c = reacting_box_case(8, 8, 8, seed=1, amp=0.0)
    mw = 28.014
    rho = p / (8314.462618 / mw * T)
    c.bv[0], c.bv[1], c.bv[2], c.bv[3], c.bv[4], c.bv[5] = rho, 30.0, -20.0, 10.0, p, T
    c.sv = np.zeros((3,) + c.bv.shape[1:])
    for s in range(3):
        c.sv[s] = y0[s]
    c.run.update(n_spec=3, n_var=8, n_scalar=3, n_scalar_transported=3, viscous_order=0, reaction=1,
                 temporal_scheme=4, fixed_time_step=1, inviscid_scheme=72)
    return c, rho


def substep_matrix(method, h):
    K = np.array([[-K1, 0, 0], [K1, -K2, 0], [0, K2, 0]])
    I = np.eye(3)
    if method == 0:
        return I + h * K
    if method == 1:
        return I + h * np.linalg.solve(I - h * K, K)
    hK = h * K
    return I + hK + hK @ hK / 2 + hK @ hK @ hK / 6


@pytest.mark.parametrize("method", [0, 1, 3])
def test_uniform_state_step_is_the_closed_form_reaction_substep(method):
    y0 = np.array([0.6, 0.3, 0.1])
    c, rho = uniform_chain_case(y0)
    c.run.update(chem_advance_method=method)
    o = Oracle(c, chain_tables())
    o.initialize()
    dt = 1.0e-7            # dt k1 = 2.5: h k1 = 0.83 per sub-step keeps the explicit update positive
    o.a["rts"][:] = dt / 2.5
    g = c.ngg
    inner = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    ry = rho * y0
    cv0 = o.a["cv"][:5][inner].copy()
    for step, (n, h) in enumerate(((3, dt / 3), (3, dt / 3))):
        o.step(dt, True)
        ry = np.linalg.matrix_power(substep_matrix(method, h), n) @ ry
        got = o.a["cv"][5:8][inner].reshape(3, -1)
        assert np.abs(got - ry[:, None]).max() <= 1e-12 * rho, (step, np.abs(got - ry[:, None]).max() / rho)
        # transport is zero on the uniform state: rho, momentum, energy unchanged
        assert np.abs(o.a["cv"][:5][inner] - cv0).max() <= 1e-12 * np.abs(cv0).max()
        ts = o.a["rts"][g:-g, g:-g, g:-g]
        assert np.allclose(ts, 1.0 / K1, rtol=1e-12)


def test_split_step_converges_to_rk3():
    tb = h2_tables()
    t_end = 4.0e-7

    def run(scheme, dt):
        c = reacting_box_case(10, 8, 8, seed=5, T_range=(1100.0, 1900.0), u_rms=30.0)
        c.run.update(viscous_order=2, fixed_time_step=1, temporal_scheme=scheme)
        o = Oracle(c, tb)
        o.initialize()
        for _ in range(int(round(t_end / dt))):
            o.step(dt, True)
        return o.a["cv"][:, 4:-4, 4:-4, 4:-4].copy()

    ref = run(3, 2.5e-9)
    errs = []
    # (at 1e-7 s the first step's explicit sub-steps of 0.2 dt / 1.2 are beyond the radicals' stability)
    for dt in (5e-8, 2.5e-8, 1.25e-8):
        d = run(4, dt)
        scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
        errs.append((np.abs(d - ref).reshape(ref.shape[0], -1).max(axis=1) / scale).max())
    ratios = np.array(errs[:-1]) / np.array(errs[1:])
    assert errs[0] < 0.1 and np.all(ratios > 1.6), (errs, ratios)
