"""Multi-block parity of the HIP path with the oracle: twelve blocks with
1-to-1 interfaces, half of them with rotated local axes (interfaces join
different local faces and running directions).  Tolerances as for one
block: normwise relative L-inf per variable <= 1e-11 per stage, <= 1e-9
after many steps."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, load_tables, multiblock_case
from oracle import Oracle

pytestmark = pytest.mark.gpu

STAGE_TOL = 1e-11
LONG_TOL = 1e-9
FIELDS = [("cv", _abi.F_CV), ("bv", _abi.F_BV), ("sv", _abi.F_SV)]


def normwise(a, b, vector=None):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    den = np.abs(b).max(axis=1)
    if vector is not None:
        den[list(vector)] = den[list(vector)].max()
    den = np.maximum(den, 1e-300)
    return np.abs(a - b).max(axis=1) / den


def check_blocks(s, o, blocks, fields, tol, info):
    for b, case in enumerate(blocks):
        for name, fid in fields:
            if name == "sv" and not case.run["n_scalar"]:
                continue
            gpu = s.download(b, fid)
            ref = o.blocks[b].a[name]
            err = normwise(gpu, ref, vector=(1, 2, 3) if name in ("cv", "bv") else None)
            assert np.all(err <= tol), f"block {b} {name}{info}: {err} > {tol}"


def _pair(blocks, tables, oracle_tables):
    mix = blocks[0].run["mixture"]
    s = Solver(blocks[0].run, tables if mix else None)
    for c in blocks:
        print(f"block {c.mx}x{c.my}x{c.mz}, {len(c.inner)} inner faces", flush=True)
        s.add_block(c)
    o = Oracle(blocks, oracle_tables if mix else None)
    return s, o


@pytest.fixture(scope="module")
def tables():
    return load_tables()


def test_multiblock_air_weno5_parity(tables):
    blocks = multiblock_case(split=(3, 2, 2), seed=7, physics="air", rotate=True)
    s, o = _pair(blocks, tables, None)
    s.initialize()
    o.initialize()
    check_blocks(s, o, blocks, FIELDS, STAGE_TOL, " after initialize")
    dt = o.cfl_dt()
    s.op(_abi.OP_COPY_QN)
    o.op("copy_qn")
    for rk in range(3):
        s.stage(rk, dt)
        o.stage(rk, dt)
        check_blocks(s, o, blocks, FIELDS, STAGE_TOL, f" after stage {rk}")
    for _ in range(20):
        dg = s.step()
        do = o.step(fixed=False)
        assert abs(dg - do) <= 1e-12 * do
    check_blocks(s, o, blocks, FIELDS, LONG_TOL, " after 21 steps")
    s.close()


def test_multiblock_h2_viscous_reacting_parity(tables, h2_oracle_tables):
    blocks = multiblock_case(split=(3, 2, 2), widths=[[9, 10, 9], [9, 10], [10, 9]], seed=5, physics="h2",
                             rotate=True)
    s, o = _pair(blocks, tables, h2_oracle_tables)
    s.initialize()
    o.initialize()
    check_blocks(s, o, blocks, FIELDS, STAGE_TOL, " after initialize")
    # the viscous operator alone (fluxes exchanged across rotated interfaces)
    s.op(_abi.OP_ZERO_DQ)
    o.op("zero_dq")
    s.op(_abi.OP_VISCOUS)
    o.op("viscous")
    for b in range(len(blocks)):
        err = normwise(s.download(b, _abi.F_DQ), o.blocks[b].a["dq"])
        assert np.all(err <= STAGE_TOL), f"block {b} viscous dq {err}"
    dt = 0.5 * o.cfl_dt()
    s.op(_abi.OP_COPY_QN)
    o.op("copy_qn")
    for rk in range(3):
        s.stage(rk, dt)
        o.stage(rk, dt)
        check_blocks(s, o, blocks, FIELDS, STAGE_TOL, f" after stage {rk}")
    for _ in range(5):
        s.step(dt)
        o.step(dt, fixed=True)
    check_blocks(s, o, blocks, FIELDS, LONG_TOL, " after 6 steps")
    for b, c in enumerate(blocks):
        g = c.ngg
        y = s.download(b, _abi.F_SV)[:, g:g + c.mz, g:g + c.my, g:g + c.mx]
        assert y.min() >= 0.0
        assert np.abs(y.sum(axis=0) - 1.0).max() <= 1e-12
    s.close()


def test_inner_face_outside_target_is_rejected(tables):
    """A face whose map leaves the target block is an error, not a launch."""
    blocks = multiblock_case(split=(2, 1, 1), widths=[[9, 9], [9], [9]], seed=1, physics="air", rotate=False)
    src, tar, tblk = blocks[0].inner[0]
    assert src[0] == src[1] and tar[4] == 1 and tar[5] > 1  # i+ face of block 0; k is the unmarked tangential axis
    bad = list(tar)
    bad[4], bad[5] = tar[4] + 5, tar[5] + 5   # the k range now starts 5 nodes in and runs past the target
    blocks[0].inner[0] = (src, bad, tblk)
    s = Solver(blocks[0].run, None)
    for c in blocks:
        s.add_block(c)
    with pytest.raises(Exception):
        s.initialize()
    s.close()
