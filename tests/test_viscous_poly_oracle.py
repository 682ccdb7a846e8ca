"""The oracle's viscous stencils pinned by polynomial exactness on a Cartesian
grid (constant metrics, constant mu, kappa):

8th order (d_dXi/d_dEta/d_dZeta<8> and the collocated fluxes,
src/ViscousScheme.cu:140-188, 470-585, 790-834), walls on every face so the
one-sided closures are used:
* the nodal fluxes (mu times first derivatives) are exact for polynomial
  fields of degree <= 4 at every node (4th-order closures at nodes 0, 1, 2),
  of degree <= 6 from node 3 inwards (6th-order central) and of degree <= 8
  from node 4 inwards (8th-order central);
* the divergence is exact for degree <= 4 everywhere and for degree <= 8 at
  nodes whose stencils of stencils stay on the 8th-order formula (8 or more
  nodes from every wall): momentum J mu (lap u + grad(div u) / 3) and energy
  J kappa lap T;
* the boundary closure is exactly 4th order: degree 5 is not exact at the wall
  node (so the checks above are not met by any stencil).

2nd order (compute_{f,g,h}v_2nd_order + compute_d{F,G,H}v,
src/ViscousScheme.cu:86-138, 836-1512): a general quadratic velocity field,
cross terms included, gives the exact momentum divergence at every node."""
import numpy as np
import pytest
from numpy.polynomial import Polynomial as P

from corefl_amd.cases import face_list, multiblock_case, reacting_box_case
from oracle import Oracle, h2_tables

MU, KAPPA = 1.7e-5, 0.031


def _poly(rng, deg):
    # random polynomial of the scaled coordinate s = x / L - 1/2 (coefficients O(1))
    return P(rng.uniform(-1, 1, deg + 1))


class Field:
    """u = A(x) + B(y) + C(z) + c1 x y, v = D(x) + E(y) + c2 y z,
    w = F(z) + G(x) + c3 x z, T = T0 (1 + 0.1 (Tx(x) + Ty(y) + Tz(z)))."""

    def __init__(self, rng, deg, L, vel=True, temp=False):
        self.L = L
        self.p = {k: _poly(rng, deg) for k in "ABCDEFG"}
        self.c = rng.uniform(-1, 1, 3) if (vel and deg >= 2) else np.zeros(3)
        if not vel:
            for k in self.p:
                self.p[k] = P([0.0])
        self.t = {k: (_poly(rng, deg) if temp else P([0.0])) for k in "xyz"}
        self.U, self.T0 = 30.0, 1000.0

    def s(self, a, x):
        return x / self.L[a] - 0.5

    def ev(self, key, a, x, der=0):
        """d^der/dx^der of polynomial key in coordinate a."""
        poly = self.p[key] if key in self.p else self.t[key]
        return poly.deriv(der)(self.s(a, x)) / self.L[a] ** der if der else poly(self.s(a, x))

    def values(self, X):
        x, y, z = X
        c1, c2, c3 = self.c
        U = self.U
        u = U * (self.ev("A", 0, x) + self.ev("B", 1, y) + self.ev("C", 2, z)) + U * c1 * x * y / (self.L[0] * self.L[1])
        v = U * (self.ev("D", 0, x) + self.ev("E", 1, y)) + U * c2 * y * z / (self.L[1] * self.L[2])
        w = U * (self.ev("F", 2, z) + self.ev("G", 0, x)) + U * c3 * x * z / (self.L[0] * self.L[2])
        T = self.T0 * (1 + 0.1 * (self.ev("x", 0, x) + self.ev("y", 1, y) + self.ev("z", 2, z)))
        return u, v, w, T

    def gradients(self, X):
        x, y, z = X
        c1, c2, c3 = self.c
        U, L = self.U, self.L
        ux = U * (self.ev("A", 0, x, 1) + c1 * y / (L[0] * L[1]))
        uy = U * (self.ev("B", 1, y, 1) + c1 * x / (L[0] * L[1]))
        uz = U * self.ev("C", 2, z, 1)
        vx = U * self.ev("D", 0, x, 1)
        vy = U * (self.ev("E", 1, y, 1) + c2 * z / (L[1] * L[2]))
        vz = U * c2 * y / (L[1] * L[2])
        wx = U * (self.ev("G", 0, x, 1) + c3 * z / (L[0] * L[2]))
        wy = 0 * x
        wz = U * (self.ev("F", 2, z, 1) + c3 * x / (L[0] * L[2]))
        Tg = [self.T0 * 0.1 * self.ev(k, a, X[a], 1) for a, k in enumerate("xyz")]
        return (ux, uy, uz, vx, vy, vz, wx, wy, wz), Tg

    def divergences(self, X):
        """div tau / mu for the three momentum rows and lap T."""
        x, y, z = X
        c1, c2, c3 = self.c
        U, L = self.U, self.L
        A2, B2, C2 = (self.ev(k, a, X[a], 2) for k, a in (("A", 0), ("B", 1), ("C", 2)))
        D2, E2 = self.ev("D", 0, x, 2), self.ev("E", 1, y, 2)
        F2, G2 = self.ev("F", 2, z, 2), self.ev("G", 0, x, 2)
        m1 = U * (4 / 3 * A2 + B2 + C2 + c3 / (3 * L[0] * L[2]))
        m2 = U * (D2 + 4 / 3 * E2 + c1 / (3 * L[0] * L[1]))
        m3 = U * (G2 + 4 / 3 * F2 + c2 / (3 * L[1] * L[2]))
        lapT = self.T0 * 0.1 * sum(self.ev(k, a, X[a], 2) for a, k in enumerate("xyz"))
        return m1, m2, m3, lapT


def _run(n, field_fn, tables):
    c = reacting_box_case(n, n, n, seed=5, amp=0.0)
    c.boundaries = face_list(n, n, n, {k: 2 for k in ("i-", "i+", "j-", "j+", "k-", "k+")})
    c.run.update(reaction=0)
    X = c.x
    L = [X[a].max() - X[a].min() for a in range(3)]
    f = field_fn(L)
    u, v, w, T = f.values(X)
    c.bv[1], c.bv[2], c.bv[3], c.bv[5] = u, v, w, T
    g = c.ngg
    for s in range(c.sv.shape[0]):
        c.sv[s] = c.sv[s][g, g, g]
    c.bv[0] = c.bv[0][g, g, g]
    c.bv[4] = c.bv[4][g, g, g]
    o = Oracle(c, tables)
    o.a["mul"][:] = MU
    o.a["kappa"][:] = KAPPA
    o.a["rhoD"][:] = 1e-5
    o.op("zero_dq")
    o.op("viscous")
    return c, o, f


def _interior(c, a):
    g = c.ngg
    return a[..., g:g + c.mz, g:g + c.my, g:g + c.mx]


def _depth(n):
    """distance of every interior node from the nearest wall (z, y, x order)"""
    d1 = np.minimum(np.arange(n), np.arange(n)[::-1])
    return np.minimum(np.minimum(d1[:, None, None], d1[None, :, None]), d1[None, None, :])


def _nodal_errors(c, o, f):
    """relative error of the nine nodal momentum fluxes and the three
    conduction fluxes against the exact stresses and kappa grad T"""
    X = _interior(c, c.x)
    (ux, uy, uz, vx, vy, vz, wx, wy, wz), Tg = f.gradients(X)
    div = ux + vy + wz
    txx, tyy, tzz = MU * (2 * ux - 2 / 3 * div), MU * (2 * vy - 2 / 3 * div), MU * (2 * wz - 2 / 3 * div)
    txy, txz, tyz = MU * (uy + vx), MU * (uz + wx), MU * (vz + wy)
    m = _interior(c, o.a["metric"])
    d = [m[0], m[4], m[8]]  # xi_x, eta_y, zeta_z on the Cartesian grid
    exact = {("ff", 1): d[0] * txx, ("ff", 2): d[0] * txy, ("ff", 3): d[0] * txz,
             ("gf", 1): d[1] * txy, ("gf", 2): d[1] * tyy, ("gf", 3): d[1] * tyz,
             ("hf", 1): d[2] * txz, ("hf", 2): d[2] * tyz, ("hf", 3): d[2] * tzz}
    err = np.zeros(X[0].shape)
    scale = max(np.abs(v).max() for v in exact.values())
    for (name, row), ex in exact.items():
        err = np.maximum(err, np.abs(_interior(c, o.a[name][row]) - ex) / scale)
    return err


@pytest.mark.parametrize("deg,depth", [(4, 0), (6, 3), (8, 4)])
def test_8th_order_nodal_fluxes_exact(deg, depth):
    n = 14
    c, o, f = _run(n, lambda L: Field(np.random.default_rng(deg), deg, L), h2_tables())
    err = _nodal_errors(c, o, f)
    dep = _depth(n)
    assert err[dep >= depth].max() <= 1e-12, (deg, err[dep >= depth].max())
    if deg > 4:
        # nearer the walls the lower-order closures are not exact for this degree
        assert err[dep < depth].max() > 1e-9


def test_8th_order_closure_is_exactly_4th_order():
    n = 14
    c, o, f = _run(n, lambda L: Field(np.random.default_rng(55), 5, L), h2_tables())
    err = _nodal_errors(c, o, f)
    assert err[_depth(n) == 0].max() > 1e-9


@pytest.mark.parametrize("deg,depth,n", [(4, 0, 12), (8, 8, 18)])
def test_8th_order_divergence_exact(deg, depth, n):
    tb = h2_tables()
    c, o, f = _run(n, lambda L: Field(np.random.default_rng(10 + deg), deg, L), tb)
    X = _interior(c, c.x)
    m1, m2, m3, _ = f.divergences(X)
    J = _interior(c, o.a["jac"])
    dq = o.a["dq"]
    sel = _depth(n) >= depth
    for row, ex in ((1, m1), (2, m2), (3, m3)):
        ref = J * MU * ex
        err = np.abs(dq[row] - ref)[sel].max() / np.abs(ref).max()
        assert err <= 1e-10, (deg, row, err)
    # conduction alone: u = 0, T polynomial -> dq_E = J kappa lap T (species uniform: no diffusion)
    c, o, f = _run(n, lambda L: Field(np.random.default_rng(20 + deg), deg, L, vel=False, temp=True), tb)
    _, _, _, lapT = f.divergences(X)
    ref = J * KAPPA * lapT
    err = np.abs(o.a["dq"][4] - ref)[sel].max() / np.abs(ref).max()
    assert err <= 1e-10, (deg, err)


def test_2nd_order_quadratic_momentum_exact():
    c = multiblock_case(split=(1, 1, 1), widths=[[10], [11], [9]], seed=3, physics="air", rotate=False,
                        wiggle=0.0)[0]
    c.run.update(viscous_order=2)
    X = c.x
    L = [X[a].max() - X[a].min() for a in range(3)]
    f = Field(np.random.default_rng(77), 2, L)
    u, v, w, _ = f.values(X)
    c.bv[0] = 1.0
    c.bv[1], c.bv[2], c.bv[3] = u, v, w
    c.bv[4] = 1.0 / 1.4
    c.bv[5] = c.bv[4] * 28.9647 / (c.bv[0] * 8314.462618)
    o = Oracle(c, None)
    o.op("properties")
    o.op("zero_dq")
    o.op("viscous")
    mu = o.a["mul"][c.ngg, c.ngg, c.ngg]
    assert np.ptp(o.a["mul"]) == 0.0
    Xi = _interior(c, X)
    J = _interior(c, o.a["jac"])
    m1, m2, m3, _ = f.divergences(Xi)
    for row, ex in ((1, m1), (2, m2), (3, m3)):
        ref = J * mu * ex
        err = np.abs(o.a["dq"][row] - ref).max() / np.abs(ref).max()
        assert err <= 1e-10, (row, err)
