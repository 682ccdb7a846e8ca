"""The Mach-10 square-body O-grid case builder: positive Jacobian on the padded
extent, a closed seam (nodes 0 and mx-1 coincide, so their metrics agree),
the far boundary fully covered by inflow and outflow ranges, and a short
oracle run that stays finite and keeps the seam nodes identical."""
import numpy as np

from corefl_amd import square_ogrid_case
from oracle import Oracle, air5_tables


def test_square_ogrid_geometry():
    c = square_ogrid_case(64, 24, seed=4)
    g = c.ngg
    assert c.jac.min() > 0
    assert np.allclose(c.x[:, :, :, g], c.x[:, :, :, g + 63], rtol=0, atol=1e-15)
    assert np.allclose(c.metric[..., g], c.metric[..., g + 63], rtol=1e-9, atol=1e-9 * np.abs(c.metric).max())
    far = [b for b in c.boundaries if b[0] == 1 and b[1] == 1]
    covered = sorted(i for b in far for i in range(b[3][0], b[4][0] + 1))
    assert covered == list(range(64))
    assert {b[2] for b in far} == {5, 6}


def test_square_ogrid_oracle_run():
    c = square_ogrid_case(48, 20, seed=4)
    o = Oracle(c, air5_tables(True))
    o.initialize()
    dt = 0.5 * o.cfl_dt()
    for _ in range(5):
        o.step(dt, fixed=True)
    g = c.ngg
    bv = o.a["bv"][:, g, g:-g, g:-g]
    assert np.isfinite(o.a["cv"]).all()
    assert np.array_equal(bv[..., 0], bv[..., -1])
    y = o.a["sv"][:5, g, g:-g, g:-g]
    assert y.min() >= 0 and np.abs(y.sum(axis=0) - 1).max() <= 1e-12
