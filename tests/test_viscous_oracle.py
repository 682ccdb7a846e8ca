"""The oracle's viscous schemes pinned by what they must satisfy.

2nd order (compute_{f,g,h}v_2nd_order + compute_d{F,G,H}v, src/ViscousScheme.cu:86-1512):
on a Cartesian grid at constant temperature (constant mu) with u = c y^2, the
face differences are exact for quadratics, so the x-momentum divergence is
2 c mu J, and the energy divergence (face-averaged u times the face stress) is
mu c^2 (6 y^2 + dy^2) J: the second difference of y^4 over dy^2.

Both orders (8th: src/ViscousScheme.cu:470-834): the mixture-averaged
diffusion fluxes with the correction velocity sum to zero over the species at
every node / face (with and without the pressure-gradient term)."""
import numpy as np
import pytest

from corefl_amd.cases import multiblock_case, reacting_box_case
from oracle import Oracle


def test_second_order_couette_identity():
    c = multiblock_case(split=(1, 1, 1), widths=[[10], [12], [9]], seed=3, physics="air", rotate=False,
                        wiggle=0.0)[0]
    c.run.update(viscous_order=2)
    g = c.ngg
    X = c.x
    y = X[1]
    cc = 0.7
    c.bv[0] = 1.0
    c.bv[1] = cc * y * y
    c.bv[2] = 0.0
    c.bv[3] = 0.0
    c.bv[4] = 1.0 / 1.4
    c.bv[5] = c.bv[4] * 28.9647 / (c.bv[0] * 8314.462618)
    o = Oracle(c, None)
    o.op("properties")
    o.op("zero_dq")
    o.op("viscous")
    mu = o.a["mul"][g, g, g]
    assert np.ptp(o.a["mul"]) == 0.0
    J = o.a["jac"][g:-g, g:-g, g:-g]
    dy = X[1, 0, 1, 0] - X[1, 0, 0, 0]
    yi = y[g:-g, g:-g, g:-g]
    dq = o.a["dq"]
    assert np.allclose(dq[1], 2 * cc * mu * J, rtol=1e-9, atol=0)
    assert np.allclose(dq[4], mu * cc * cc * (6 * yi * yi + dy * dy) * J, rtol=1e-8,
                       atol=1e-10 * np.abs(dq[4]).max())
    assert np.abs(dq[2]).max() <= 1e-12 * np.abs(dq[1]).max()


@pytest.mark.parametrize("order", [2, 8])
@pytest.mark.parametrize("gradp", [0, 1])
def test_species_diffusion_fluxes_sum_to_zero(order, gradp, h2_oracle_tables):
    c = reacting_box_case(14, 12, 10, seed=31)
    c.run.update(viscous_order=order, gradp_in_diffusion=gradp, reaction=0)
    o = Oracle(c, h2_oracle_tables)
    o.initialize()
    o.op("zero_dq")
    o.op("viscous")
    g = c.ngg
    if order == 8:
        # nodal species fluxes: planes 5..13 of f/g/hFlux
        for name in ("ff", "gf", "hf"):
            f = o.a[name][5:14, g:-g, g:-g, g:-g]
            assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).max()
    else:
        # the last direction's face fluxes stay in fFlux: species components 4..12
        f = o.a["ff"][4:13, g - 1:-g, g:-g, g:-g]
        assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).max()
    # and therefore the species source terms of the divergence sum to zero too
    dq = o.a["dq"]
    assert np.abs(dq[5:14].sum(axis=0)).max() <= 1e-11 * np.abs(dq[5:14]).max()
