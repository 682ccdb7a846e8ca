"""The tile-shape autotuner of the characteristic WENO sweeps: every shape
computes the same faces with the same arithmetic, so the default shapes, the
alternative shapes on every direction and the shapes the tuner picks per block
must give the same state bit for bit (9-species WENO7 box with chemistry and
8th-order viscous terms; twelve Air WENO5 blocks of different sizes)."""
import numpy as np
import pytest

from corefl_amd import Solver, _abi, load_tables, multiblock_case, reacting_box_case

pytestmark = pytest.mark.gpu


def run(cases, tables, mode, steps=3):
    r = dict(cases[0].run)
    r["autotune"] = 1 if mode == "tuned" else 0
    s = Solver(r, tables)
    for c in cases:
        s.add_block(c)
    if mode == "alternative":
        for b in range(len(cases)):
            s.set_tiles(b, [1, 1, 1])
    s.initialize()
    dt = 0.4 * s.cfl_dt()
    for _ in range(steps):
        s.step(dt)
    out = [[s.download(b, f) for f in (_abi.F_CV, _abi.F_BV, _abi.F_SV)] for b in range(len(cases))]
    tiles = [s.tiles(b) for b in range(len(cases))]
    s.close()
    return out, tiles


@pytest.mark.parametrize("physics", ["h2", "air"])
def test_tile_shapes_give_identical_results(physics):
    if physics == "h2":
        tables = load_tables()
        cases = [reacting_box_case(40, 24, 20, seed=51, tables=tables)]
    else:
        tables = None
        cases = multiblock_case(split=(3, 2, 2), seed=13, physics="air", rotate=True)
    print(f"autotune {physics}: {len(cases)} block(s)", flush=True)
    ref, t0 = run(cases, tables, "default")
    alt, t1 = run(cases, tables, "alternative")
    tuned, t2 = run(cases, tables, "tuned")
    assert all(t == [0, 0, 0] for t in t0) and all(t == [1, 1, 1] for t in t1)
    assert all(v in (0, 1) for t in t2 for v in t)
    for b in range(len(cases)):
        for x, y, z in zip(ref[b], alt[b], tuned[b]):
            assert np.array_equal(x, y), f"block {b}: alternative shapes differ"
            assert np.array_equal(x, z), f"block {b}: tuned shapes {t2[b]} differ"
