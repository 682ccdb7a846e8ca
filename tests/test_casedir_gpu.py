"""Reference cases run from their case directories on the GPU, read by the
C++ readers (corefl_setup_load, corefl_mesh_*):

* sodShockTube: the example's setup.txt as shipped (WENO7 characteristic,
  SSP-RK3, fixed dt 1e-4, outflow faces, the two-state initial groups) with
  the grid its generate1DMesh.py describes, written in readGrid's format.
  Fifty steps match the oracle's run of the same Case to 1e-9 normwise; the
  GPU run continued to the setup's total_simulation_time 0.2 lies on the
  reference's exact solution (example/sodShockTube/exact.dat): mean density
  and velocity errors below 5e-3 / 8e-3 on the centre line (the oracle's
  own run of the case: 3.5e-3 / 5.8e-3 on these 201 nodes)."""
from pathlib import Path

import numpy as np
import pytest

from corefl_amd import Solver, _abi
from corefl_amd.casedir import load_case_dir
from oracle import Oracle
from test_casedir import make_sod_dir

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_sod_case_dir_runs_to_the_exact_solution(tmp_path):
    cd = load_case_dir(make_sod_dir(tmp_path))
    case = cd.cases[0]
    s = Solver(case.run, None)
    s.add_block(case)
    s.initialize()
    o = Oracle(case, None)
    o.initialize()
    dt = case.run["dt"]
    print("sodShockTube case directory: 50 steps against the oracle", flush=True)
    for _ in range(50):
        s.step(dt)
        o.step(dt, True)
    g = case.ngg
    cv = s.download(0, _abi.F_CV).reshape(o.a["cv"].shape)
    a, r = cv[:, g, g:-g, g:-g].reshape(5, -1), o.a["cv"][:, g, g:-g, g:-g].reshape(5, -1)
    err = np.abs(a - r).max(axis=1) / np.abs(r).max(axis=1).clip(1e-300)
    err[1] = np.abs(a[1:4] - r[1:4]).max() / np.abs(r[1:4]).max()
    assert np.all(err[[0, 1, 4]] <= 1e-9), err
    t_end = 0.2
    steps = int(round(t_end / dt)) - 50
    print(f"sodShockTube: {steps} more steps to t = {t_end}", flush=True)
    for _ in range(steps):
        s.step(dt)
    bv = s.download(0, _abi.F_BV).reshape(o.a["bv"].shape)
    rho = bv[0, g, g + 3, g:g + case.mx]
    u = bv[1, g, g + 3, g:g + case.mx]
    x = case.x[0, g, g + 3, g:g + case.mx]
    ex = np.array(open(GOLDEN / "sod" / "exact.dat").read().split()[1:], dtype=float).reshape(-1, 4)
    assert np.abs(rho - np.interp(x, ex[:, 0], ex[:, 2])).mean() < 5e-3
    assert np.abs(u - np.interp(x, ex[:, 0], ex[:, 1])).mean() < 8e-3
    s.close()
    cd.close()


def test_square_cylinder_case_dir_matches_the_oracle(tmp_path):
    """squareCylinder2D_M10_air5_dns from its setup.txt, boundary and inner
    files as shipped (24 blocks, two-temperature AIR-5 with NASA-9, WENO7,
    8th-order viscous terms, V-T relaxation, positivity limiter, limit_flow,
    Mach-10 freestream onto a quiescent field, adiabatic wall) on a grid of the
    example's block structure: three steps of the setup's dt = 5e-10 s match
    the oracle's run of the same 24 Cases to 1e-9 normwise per variable over
    all blocks (momentum as one vector)."""
    from test_casedir import make_square_cylinder_dir
    from oracle import air5_tables
    cd = load_case_dir(make_square_cylinder_dir(tmp_path))
    s = Solver(cd.run, cd.tables)
    for c in cd.cases:
        s.add_block(c)
    s.initialize()
    o = Oracle(cd.cases, air5_tables(True))
    o.initialize()
    dt = cd.run["dt"]
    print(f"squareCylinder2D case directory: {len(cd.cases)} blocks, 3 steps of {dt}", flush=True)
    for _ in range(3):
        s.step(dt)
        o.step(dt, True)
    # normwise per variable over the whole field (most blocks are still quiescent after three steps: their
    # momentum is round-off, so a per-block scale would compare round-off with round-off)
    got, ref = [], []
    for b, blk in enumerate(o.blocks):
        g = cd.cases[b].ngg
        cv = s.download(b, _abi.F_CV).reshape(blk.a["cv"].shape)
        got.append(cv[:, g, g:-g, g:-g].reshape(cv.shape[0], -1))
        ref.append(blk.a["cv"][:, g, g:-g, g:-g].reshape(cv.shape[0], -1))
    a, r = np.concatenate(got, axis=1), np.concatenate(ref, axis=1)
    scale = np.abs(r).max(axis=1)
    scale[1:4] = scale[1:4].max()
    err = np.abs(a - r).max(axis=1) / scale
    print("normwise errors per variable:", np.array2string(err, precision=2), flush=True)
    worst = err.max()
    assert worst <= 1e-9, worst
    # the freestream entered: the inflow faces' neighbourhood moves
    assert np.abs(s.download(0, _abi.F_BV).reshape(o.blocks[0].a["bv"].shape)[1]).max() > 100.0
    s.close()
    cd.close()
