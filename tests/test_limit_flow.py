"""limit_flow (src/TimeAdvanceFunc.cuh:129-232): an interior point with a NaN
or an out-of-bounds rho, u, v, w or p (bounds from the reference state,
src/DParameter.cu:371-387) takes the mean of the good points of its 3x3x3
neighbourhood, T from p / (rho R), cv from the primitives; every repair is
computed from the state before any repair.  Oracle pins on the CPU, parity
of the HIP path on the GPU."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, load_tables, reacting_box_case
from corefl_amd.cases import multiblock_case
from oracle import Oracle

R_AIR = 8314.462618 / 28.9647


def planted(physics, tables=None):
    if physics == "air":
        c = multiblock_case(split=(1, 1, 1), widths=[[12], [10], [9]], seed=4, physics="air", rotate=False)[0]
        c.run.update(limit_flow=1, p_ref=1.0)
    else:
        c = reacting_box_case(12, 10, 9, seed=8, tables=tables)
        c.run.update(limit_flow=1, p_ref=101325.0)
    g = c.ngg
    c.bv[0, g + 4, g + 5, g + 6] = np.nan                       # lone NaN
    c.bv[4, g + 2, g + 3, g + 3] = -1.0                         # negative pressure ...
    c.bv[1, g + 2, g + 3, g + 4] = 1e7 * c.run["v_ref"]         # ... next to a runaway velocity
    c.bv[0, g + 0, g + 0, g + 0] = 1e9 * c.run["rho_ref"]       # corner point
    return c


def test_oracle_repairs_from_good_neighbours():
    c = planted("air")
    g = c.ngg
    before = c.bv.copy()
    o = Oracle(c, None)
    o.op("cv_from_bv")
    o.op("limit_flow")
    bv = o.a["bv"]
    bad = {(4, 5, 6), (2, 3, 3), (2, 3, 4), (0, 0, 0)}
    for (k, j, i) in bad:
        acc, n = np.zeros(5), 0
        for kk in range(max(k - 1, 0), min(k + 2, c.mz)):
            for jj in range(max(j - 1, 0), min(j + 2, c.my)):
                for ii in range(max(i - 1, 0), min(i + 2, c.mx)):
                    if (kk, jj, ii) in bad:
                        continue
                    acc += before[:5, g + kk, g + jj, g + ii]
                    n += 1
        ref = acc / n
        got = bv[:5, g + k, g + j, g + i]
        assert np.allclose(got, ref, rtol=1e-14, atol=0)
        assert abs(bv[5, g + k, g + j, g + i] - ref[4] / (ref[0] * R_AIR)) <= 1e-14 * bv[5, g + k, g + j, g + i]
    # every other point is untouched
    mask = np.ones(bv.shape[1:], bool)
    for (k, j, i) in bad:
        mask[g + k, g + j, g + i] = False
    assert np.array_equal(np.nan_to_num(bv[:, mask]), np.nan_to_num(before[:, mask]))


def test_oracle_clips_when_no_neighbour_is_good():
    c = multiblock_case(split=(1, 1, 1), widths=[[8], [8], [8]], seed=4, physics="air", rotate=False)[0]
    c.run.update(limit_flow=1, p_ref=1.0)
    g = c.ngg
    c.bv[4, g:g + 3, g:g + 3, g:g + 3] = 1e9      # a 3x3x3 block of bad pressure: its centre has no good neighbour
    o = Oracle(c, None)
    o.op("cv_from_bv")
    o.op("limit_flow")
    assert o.a["bv"][4, g + 1, g + 1, g + 1] == 1e3 * 1.0


@pytest.mark.gpu
@pytest.mark.parametrize("physics", ["air", "h2"])
def test_limit_flow_parity(physics, h2_oracle_tables):
    tables = load_tables()
    c = planted(physics, tables)
    mix = c.run["mixture"]
    s = Solver(c.run, tables if mix else None)
    s.add_block(c)
    o = Oracle(c, h2_oracle_tables if mix else None)
    s.op(_abi.OP_CV_FROM_BV)
    o.op("cv_from_bv")
    s.op(_abi.OP_LIMIT_FLOW)
    o.op("limit_flow")
    for name, fid in (("bv", _abi.F_BV), ("cv", _abi.F_CV)):
        a, b = s.download(0, fid), o.a[name]
        assert np.array_equal(np.isnan(a), np.isnan(b))
        a, b = np.nan_to_num(a), np.nan_to_num(b)
        err = np.abs(a - b).reshape(a.shape[0], -1).max(axis=1) / np.maximum(np.abs(b).reshape(b.shape[0], -1).max(axis=1), 1e-300)
        assert np.all(err <= 1e-13), f"{name}: {err}"
    s.close()
