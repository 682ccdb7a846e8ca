"""Reference case directories read by the C++ readers (corefl_setup_load,
corefl_mesh_load; src/Mesh.cpp:692-962, 1013-1217, 244-528): the shipped
sodShockTube setup (tests/golden/cases/sodShockTube/input/setup.txt, the
example's file) with the grid its generate1DMesh.py describes (201 x 7 nodes
on [0, 1] x [-0.03, 0.03], outflow faces) written here in readGrid's binary
and text formats.  No GPU: the readers are host code.

* the coordinates read back equal the written ones, binary and text alike;
* the ghost nodes continue the uniform lattice (mirror extension), so the
  metrics are 1/dx, 1/dy on the diagonal and the Jacobian dx dy (unit z
  spacing in 2-D) over the whole padded extent, as the central formula of
  metrics() gives on the same padded coordinates;
* the grouped initial state puts the right state on the nodes inside
  init_cond_0's box (x in [0.5, 1.5], z in [-2, 2]) and the left state
  elsewhere, ghost nodes included, with T = p MW / (rho R);
* the four faces carry label 6 and the outflow condition of the setup;
* ngg is 4 (WENO7, inviscid_scheme 72)."""
import shutil
from pathlib import Path

import numpy as np
import pytest

from corefl_amd.casedir import load_case_dir, write_boundary, write_grid
from corefl_amd.cases import metrics

GOLDEN = Path(__file__).resolve().parent / "golden" / "cases"
R_AIR = 8314.462618 / 28.9647


def sod_grid(nx=201, ny=7):
    x = np.linspace(0.0, 1.0, nx)
    y = np.linspace(-0.03, 0.03, ny)
    X = np.broadcast_to(x[None, None, :], (1, ny, nx)).copy()
    Y = np.broadcast_to(y[None, :, None], (1, ny, nx)).copy()
    Z = np.zeros_like(X)
    faces = [(0, 0, 0, ny - 1, 0, 0, 6), (nx - 1, nx - 1, 0, ny - 1, 0, 0, 6), (0, nx - 1, 0, 0, 0, 0, 6),
             (0, nx - 1, ny - 1, ny - 1, 0, 0, 6)]
    return (X, Y, Z), faces


def make_sod_dir(root: Path, binary: bool = True) -> Path:
    case = root / "sodShockTube"
    (case / "input").mkdir(parents=True)
    text = (GOLDEN / "sodShockTube" / "input" / "setup.txt").read_text()
    if not binary:
        text = text.replace("int gridIsBinary = 1", "int gridIsBinary = 0")
    (case / "input" / "setup.txt").write_text(text)
    xyz, faces = sod_grid()
    write_grid(case / "input", [xyz], binary=binary)
    write_boundary(case / "input", [faces])
    return case


@pytest.mark.parametrize("binary", [True, False])
def test_sod_case_dir(tmp_path, binary):
    cd = load_case_dir(make_sod_dir(tmp_path, binary))
    assert len(cd.cases) == 1
    c = cd.cases[0]
    assert (c.mx, c.my, c.mz, c.ngg) == (201, 7, 1, 4)
    g = c.ngg
    (X, Y, _), _ = sod_grid()
    assert np.array_equal(c.x[0][g, g:-g, g:-g], X[0]) and np.array_equal(c.x[1][g, g:-g, g:-g], Y[0])
    dx, dy = 1.0 / 200, 0.01
    # ghost coordinates continue the lattice
    assert np.allclose(np.diff(c.x[0], axis=2), dx, rtol=1e-9) and np.allclose(np.diff(c.x[1], axis=1), dy, rtol=1e-9)
    assert np.allclose(np.diff(c.x[2], axis=0), 1.0)
    assert np.allclose(c.metric[0], 1 / dx, rtol=1e-9) and np.allclose(c.metric[4], 1 / dy, rtol=1e-9)
    assert np.allclose(c.metric[8], 1.0, rtol=1e-12)
    for l in (1, 2, 3, 5, 6, 7):
        assert np.abs(c.metric[l]).max() <= 1e-9
    assert np.allclose(c.jac, dx * dy, rtol=1e-9)
    # the same central formula on the padded coordinates (extended by one more node)
    pad = lambda A, ax, h: np.concatenate([np.take(A, [0], ax) - h, A, np.take(A, [-1], ax) + h], axis=ax)
    Xp, Yp, Zp = c.x[0], c.x[1], c.x[2]
    for ax, h in ((2, (dx, 0, 0)), (1, (0, dy, 0)), (0, (0, 0, 1.0))):
        Xp, Yp, Zp = pad(Xp, ax, h[0]), pad(Yp, ax, h[1]), pad(Zp, ax, h[2])
    m_ref, j_ref = metrics(Xp, Yp, Zp)
    assert np.allclose(c.metric, m_ref, rtol=1e-9, atol=1e-9) and np.allclose(c.jac, j_ref, rtol=1e-9)
    # grouped initial state
    # (the 2-D ghost planes sit at z = +-1 .. +-4: those beyond the box's z range take the default state)
    right = (c.x[0] >= 0.5) & (c.x[0] <= 1.5) & (np.abs(c.x[2]) <= 2.0)
    assert np.all(c.bv[0][right] == 0.125) and np.all(c.bv[0][~right] == 1.0)
    assert np.all(c.bv[4][right] == 0.1) and np.all(c.bv[4][~right] == 1.0)
    assert np.allclose(c.bv[5], c.bv[4] / (c.bv[0] * R_AIR), rtol=1e-14)
    assert np.all(c.bv[1:4] == 0.0)
    # faces and conditions
    assert sorted((f[0], f[1], f[2]) for f in c.boundaries) == [(0, -1, 6), (0, 1, 6), (1, -1, 6), (1, 1, 6)]
    assert [(d["label"], d["kind"]) for d in cd.conditions] == [(6, 6)]
    assert c.run["inviscid_scheme"] == 72 and c.run["fixed_time_step"] == 1 and c.run["dt"] == 1e-4
    cd.close()


DATA = Path(__file__).resolve().parent.parent / "corefl-cpc_amd" / "data" / "air5_park"
MB = Path(__file__).resolve().parent / "golden" / "multiblock"
# the square cylinder's 24 blocks: node counts of the five columns and rows (the body takes the middle)
SQ_COLS, SQ_ROWS = [89, 73, 133, 73, 221], [89, 73, 133, 73, 89]


def square_cylinder_grid(side=0.01):
    """Blocks in the file's order (row-major, the body's block left out) on a
    uniform lattice of spacing side / 132 with the body's lower-left corner at
    the origin; the example's own grid comes from its generator, which is
    not run here, so the spacing is ours and the block structure theirs."""
    h = side / 132
    c0 = np.concatenate([[0], np.cumsum([n - 1 for n in SQ_COLS])])
    r0 = np.concatenate([[0], np.cumsum([n - 1 for n in SQ_ROWS])])
    blocks = []
    for r in range(5):
        for c in range(5):
            if r == 2 and c == 2:
                continue
            x = (c0[c] + np.arange(SQ_COLS[c]) - c0[2]) * h
            y = (r0[r] + np.arange(SQ_ROWS[r]) - r0[2]) * h
            X = np.broadcast_to(x[None, None, :], (1, len(y), len(x))).copy()
            Y = np.broadcast_to(y[None, :, None], (1, len(y), len(x))).copy()
            blocks.append((X, Y, np.zeros_like(X)))
    return blocks


def make_square_cylinder_dir(root: Path) -> Path:
    case = root / "squareCylinder2D"
    inp = case / "input"
    (inp / "chemistry").mkdir(parents=True)
    (inp / "boundary_condition").mkdir()
    shutil.copy(GOLDEN / "squareCylinder2D" / "input" / "setup.txt", inp / "setup.txt")
    for src, dst in (("mechanism.inp", "air5_park.inp"), ("therm.dat", "therm.dat"), ("tran.dat", "tran.dat"),
                     ("two_temperature.dat", "two_temperature.dat")):
        shutil.copy(DATA / src, inp / "chemistry" / dst)
    shutil.copy(MB / "square_cylinder_boundary.txt", inp / "boundary_condition" / "boundary   0.txt")
    shutil.copy(MB / "square_cylinder_inner.txt", inp / "boundary_condition" / "inner   0.txt")
    write_grid(inp, square_cylinder_grid())
    return case


def test_square_cylinder_case_dir(tmp_path):
    """The example's setup, boundary and inner files: 24 blocks, ngg 4, the
    freestream / outflow / symmetry / adiabatic-wall conditions, the quiescent
    initial state, interface ghost nodes that continue the neighbour's
    lattice (so every metric is the uniform one, also across interfaces)."""
    cd = load_case_dir(make_square_cylinder_dir(tmp_path))
    assert len(cd.cases) == 24 and cd.run["two_temperature"] == 1 and cd.run["n_spec"] == 5
    h = 0.01 / 132
    for c in cd.cases:
        assert c.ngg == 4 and c.mz == 1
        assert np.allclose(c.metric[0], 1 / h, rtol=1e-7) and np.allclose(c.metric[4], 1 / h, rtol=1e-7)
        assert np.all(c.bv[1:4] == 0.0) and np.allclose(c.bv[5], 300.0) and np.allclose(c.bv[4], 1000.0)
        assert np.allclose(c.tve, 300.0)
    kinds = {d["label"]: d for d in cd.conditions}
    assert kinds[5]["kind"] == 5 and kinds[6]["kind"] == 6 and kinds[3]["kind"] == 3
    assert kinds[2]["kind"] == 2 and kinds[2]["thermal_type"] == 1
    fs = kinds[5]
    R = 8314.462618 / fs["mw"]
    assert abs(fs["density"] - 1000.0 / (R * 300.0)) <= 1e-12 * fs["density"]
    assert fs["u"] > 3000 and fs["v"] == 0.0  # Mach 10 at 300 K
    cd.close()
