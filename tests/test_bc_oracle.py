"""Boundary conditions of the oracle pinned by what each condition imposes
(apply_symmetry src/BoundCond.cu:2130-2202, apply_inflow 2484-2997 constant
branch, apply_wall 3055-3387): symmetry planes carry no normal velocity and
mirror the interior; walls are no-slip with the wall (or interior)
temperature at the face and the wall temperature as the face average of
ghost and interior; a supersonic inflow imposes its state; a subsonic
characteristic inflow leaves a flow already equal to the inflow state
unchanged."""
import numpy as np
import pytest

from corefl_amd import channel_case
from oracle import Oracle


def unit_normal(case, o, face, i, j, k):
    g = case.ngg
    m = o.a["metric"][3 * face:3 * face + 3, k + g, j + g, i + g]
    return m / np.linalg.norm(m)


def at(o, case, name, i, j, k):
    g = case.ngg
    return o.a[name][:, k + g, j + g, i + g]


@pytest.fixture(params=["air", "h2"])
def applied(request, h2_oracle_tables):
    case = channel_case(24, 14, 12, physics=request.param, subsonic=False, seed=9)
    o = Oracle(case, h2_oracle_tables if request.param == "h2" else None)
    o.op("properties")
    o.op("bc")
    return case, o


def test_symmetry_planes(applied):
    case, o = applied
    for k, d in ((0, -1), (case.mz - 1, 1)):
        for j in range(2, case.my - 2, 3):
            for i in range(2, case.mx - 2, 4):
                n = unit_normal(case, o, 2, i, j, k)
                vf = at(o, case, "bv", i, j, k)[1:4]
                vin = at(o, case, "bv", i, j, k - d)[1:4]
                scale = np.abs(vin).max()
                assert abs(vf @ n) <= 1e-13 * scale
                assert np.allclose(vf, vin - (vin @ n) * n, rtol=0, atol=1e-13 * scale)
                for gl in range(1, case.ngg + 1):
                    bg = at(o, case, "bv", i, j, k + gl * d)
                    bi = at(o, case, "bv", i, j, k - gl * d)
                    assert abs((bg[1:4] + bi[1:4]) @ n) <= 1e-12 * scale
                    tg, ti = bg[1:4] - (bg[1:4] @ n) * n, bi[1:4] - (bi[1:4] @ n) * n
                    assert np.allclose(tg, ti, rtol=0, atol=1e-12 * scale)
                    assert np.array_equal(bg[[0, 4, 5]], bi[[0, 4, 5]])


def test_walls(applied):
    case, o = applied
    conds = {c["label"]: c for c in case.conditions}
    t_iso = conds[2]["wall_temperature"]
    for j, d, adiabatic in ((0, -1, False), (case.my - 1, 1, True)):
        for k in range(2, case.mz - 2, 3):
            for i in range(2, case.mx - 2, 4):
                bf = at(o, case, "bv", i, j, k)
                assert np.all(bf[1:4] == 0.0)
                t_in = at(o, case, "bv", i, j - d, k)[5]
                assert bf[5] == (t_in if adiabatic else t_iso)
                assert bf[4] == at(o, case, "bv", i, j - d, k)[4]
                for gl in range(1, case.ngg + 1):
                    bg = at(o, case, "bv", i, j + gl * d, k)
                    bi = at(o, case, "bv", i, j - gl * d, k)
                    assert np.array_equal(bg[1:4], -bi[1:4])
                    assert bg[4] == bi[4]
                    if adiabatic:
                        assert bg[5] == bi[5]
                    elif 2 * t_iso - bi[5] > 0.1 * t_iso:
                        assert abs(0.5 * (bg[5] + bi[5]) - t_iso) <= 1e-12 * t_iso
                    else:
                        assert bg[5] == t_iso


def test_supersonic_inflow_imposes_its_state(applied):
    case, o = applied
    c = {c["label"]: c for c in case.conditions}[5]
    state = np.array([c["density"], c["u"], c["v"], c["w"], c["pressure"], c["temperature"]])
    # edge nodes also belong to wall and symmetry faces, which come later in the priority order
    for k in range(1, case.mz - 1, 3):
        for j in range(1, case.my - 1, 3):
            for gl in range(0, case.ngg + 1):
                assert np.array_equal(at(o, case, "bv", -gl, j, k), state)


@pytest.mark.parametrize("physics", ["air", "h2"])
def test_subsonic_inflow_keeps_a_matching_flow(physics, h2_oracle_tables):
    case = channel_case(24, 14, 12, physics=physics, subsonic=True, seed=4, wiggle=0.0)
    c = {cd["label"]: cd for cd in case.conditions}[5]
    case.bv[:] = np.array([c["density"], c["u"], c["v"], c["w"], c["pressure"], c["temperature"]])[:, None, None,
                                                                                                      None]
    if physics == "h2":
        case.sv[:] = np.array(c["sv"])[:, None, None, None]
    o = Oracle(case, h2_oracle_tables if physics == "h2" else None)
    o.op("properties")
    o.op("bc")
    for k in range(1, case.mz - 1, 2):
        for j in range(1, case.my - 1, 2):
            for gl in range(0, case.ngg + 1):
                b = at(o, case, "bv", -gl, j, k)
                ref = np.array([c["density"], c["u"], c["v"], c["w"], c["pressure"], c["temperature"]])
                assert np.allclose(b, ref, rtol=1e-12, atol=1e-12 * abs(c["u"]))
