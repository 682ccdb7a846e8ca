"""Synthetic cases: grids, metrics, boundary lists and seeded initial fields.

Metrics follow the reference's second-order central formula
(Block::compute_jac_metric, src/Mesh.cpp:244-528) evaluated over the padded
extent from coordinates that extend one node beyond the ghost layers.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _abi

R_U = 8314.462618
MW_AIR = 28.9647
P_ATM = 101325.0
DATA = Path(__file__).resolve().parent.parent / "data"
H2_MECH = DATA / "h2_li2004"


@dataclass
class Case:
    """Everything needed to set up one block (arrays over the padded extent)."""
    mx: int
    my: int
    mz: int
    ngg: int
    metric: np.ndarray          # (9, nzp, nyp, nxp)
    jac: np.ndarray             # (nzp, nyp, nxp)
    bv: np.ndarray              # (6, nzp, nyp, nxp): rho, u, v, w, p, T
    sv: np.ndarray              # (n_scalar, nzp, nyp, nxp)
    boundaries: list = field(default_factory=list)  # (face, direction, type, start[3], end[3])
    run: dict = field(default_factory=dict)
    x: np.ndarray | None = None  # node coordinates over the padded extent (3, nzp, nyp, nxp)
    # 1-to-1 interfaces in the notation of boundary_condition/inner%4d.txt:
    # (source six numbers, target six numbers, 1-based target block)
    inner: list = field(default_factory=list)
    # local -> global lattice map of a block of a multi-block case:
    # (perm, sign, gmin, gmax) with global[perm[a]] = gmin + l_a (sign +1) or gmax - l_a (sign -1)
    lattice: tuple | None = None
    # conditions of boundary labels (the setup's structs), dicts with the
    # fields of corefl_bc_condition: label, kind, thermal_type, inflow_sub,
    # wall_temperature, density, u, v, w, pressure, temperature, mw, sv
    conditions: list = field(default_factory=list)
    # two-temperature runs: the vibrational-electronic temperature over the
    # padded extent (nzp, nyp, nxp); the state's E_ve scalar follows from it
    tve: np.ndarray | None = None

    @property
    def padded_shape(self):
        g = self.ngg
        return (self.mz + 2 * g, self.my + 2 * g, self.mx + 2 * g)

    @property
    def points(self):
        return self.mx * self.my * self.mz


def metrics(X, Y, Z):
    """Jacobian and metrics from coordinates with one extra node on every side.

    X, Y, Z have shape (nzp+2, nyp+2, nxp+2); the result covers the inner
    (nzp, nyp, nxp) nodes.  Central differences as src/Mesh.cpp:254-292.
    """
    def d(A, ax):
        sl_p = [slice(1, -1)] * 3
        sl_m = [slice(1, -1)] * 3
        sl_p[ax] = slice(2, None)
        sl_m[ax] = slice(0, -2)
        return (A[tuple(sl_p)] - A[tuple(sl_m)]) * 0.5

    # axis 2 = i (xi), 1 = j (eta), 0 = k (zeta)
    dxd1, dyd1, dzd1 = d(X, 2), d(Y, 2), d(Z, 2)
    dxd2, dyd2, dzd2 = d(X, 1), d(Y, 1), d(Z, 1)
    dxd3, dyd3, dzd3 = d(X, 0), d(Y, 0), d(Z, 0)
    jac = (dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
           dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1)
    if np.any(jac <= 0):
        raise ValueError("non-positive jacobian")
    m = np.empty((9,) + jac.shape)
    m[0] = (dyd2 * dzd3 - dyd3 * dzd2) / jac
    m[1] = (dxd3 * dzd2 - dxd2 * dzd3) / jac
    m[2] = (dxd2 * dyd3 - dxd3 * dyd2) / jac
    m[3] = (dyd3 * dzd1 - dyd1 * dzd3) / jac
    m[4] = (dxd1 * dzd3 - dxd3 * dzd1) / jac
    m[5] = (dxd3 * dyd1 - dxd1 * dyd3) / jac
    m[6] = (dyd1 * dzd2 - dyd2 * dzd1) / jac
    m[7] = (dxd2 * dzd1 - dxd1 * dzd2) / jac
    m[8] = (dxd1 * dyd2 - dxd2 * dyd1) / jac
    return np.ascontiguousarray(m), np.ascontiguousarray(jac)


def face_list(mx, my, mz, types):
    """Boundary faces in the order i-, i+, j-, j+, (k-, k+); types = dict face-key -> label."""
    out = []
    n = (mx, my, mz)
    keys = [("i-", 0, -1), ("i+", 0, 1), ("j-", 1, -1), ("j+", 1, 1)]
    if mz > 1:
        keys += [("k-", 2, -1), ("k+", 2, 1)]
    for key, face, direction in keys:
        if key not in types:
            continue
        start = [0, 0, 0]
        end = [mx - 1, my - 1, mz - 1]
        at = 0 if direction < 0 else n[face] - 1
        start[face] = end[face] = at
        out.append((face, direction, types[key], start, end))
    return out


def _grid_nodes(mx, my, mz, g, coord):
    """Coordinates over the padded extent plus one node (for metrics)."""
    kk = np.arange(-g - 1, mz + g + 1)
    jj = np.arange(-g - 1, my + g + 1)
    ii = np.arange(-g - 1, mx + g + 1)
    K, J, I = np.meshgrid(kk, jj, ii, indexing="ij")
    return coord(I.astype(float), J.astype(float), K.astype(float))


# ---------------------------------------------------------------------------
# 1-D Sod shock tube on a 2-D grid (the reference needs several rows in y)
# ---------------------------------------------------------------------------
def sod_case(nx: int = 400, ny: int = 7, scheme: int = 52, cfl: float = 0.5) -> Case:
    g = 3 if scheme == 52 else 4
    dx = 1.0 / (nx - 1)
    dy = 0.01
    dz = dx

    def coord(I, J, K):
        return I * dx, J * dy, K * dz

    X, Y, Z = _grid_nodes(nx, ny, 1, g, coord)
    metric, jac = metrics(X, Y, Z)
    shp = jac.shape
    xc = X[1:-1, 1:-1, 1:-1]
    bv = np.zeros((6,) + shp)
    right = xc >= 0.5
    bv[0] = np.where(right, 0.125, 1.0)
    bv[4] = np.where(right, 0.1, 1.0)
    bv[5] = bv[4] * MW_AIR / (bv[0] * R_U)
    sv = np.zeros((0,) + shp)
    rho_ref, p_ref = 1.0, 1.0
    c_ref = np.sqrt(1.4 * p_ref / rho_ref)
    run = dict(mixture=0, n_spec=0, n_var=5, n_scalar=0, n_scalar_transported=0, inviscid_scheme=scheme,
               viscous_order=0, reaction=0, gradp_in_diffusion=0, fixed_time_step=0, dt=0.0, cfl=cfl,
               prandtl=0.72, rho_ref=rho_ref, v_ref=c_ref)
    faces = face_list(nx, ny, 1, {"i-": 6, "i+": 6, "j-": 6, "j+": 6})
    return Case(nx, ny, 1, g, metric, jac, bv, sv, faces, run, np.stack([xc, Y[1:-1, 1:-1, 1:-1],
                                                                          Z[1:-1, 1:-1, 1:-1]]))


# ---------------------------------------------------------------------------
# One served 1-D shock tube as the one-block case the solver and the oracle run
# ---------------------------------------------------------------------------
def tube_case(n_cells: int, length: float, left, right, tables: _abi.ChemTables, n_left: int | None = None,
              bc=("outflow", "outflow"), scheme: int = 72, reaction: int = 1, cfl: float = 0.5) -> Case:
    """The tube of corefl_tubes_submit on a (n_cells, 2 g + 1, 1) block whose
    rows all hold the tube (outflow faces in j, so the middle row j = g sees
    no y-ghost in its stencils): the same lattice spacing length / (n - 1) in
    x, y and z, metrics 1 / dx on the diagonal, jac = dx^3, and the initial
    density p / (R T) with R summed in species order, all computed as the
    tube server computes them.  left / right: (T, p, u, Y)."""
    g = 4 if scheme in (71, 72) else 3
    ny = 2 * g + 1
    n_left = n_cells // 2 if n_left is None else n_left
    dx = length / (n_cells - 1)
    shp = (1 + 2 * g, ny + 2 * g, n_cells + 2 * g)
    metric = np.zeros((9,) + shp)
    inv = 1.0 / dx
    metric[0] = metric[4] = metric[8] = inv
    jac = np.full(shp, dx * dx * dx)
    ns = tables.n_spec
    gas = [float(tables.gas_const[l]) for l in range(ns)]
    bv = np.zeros((6,) + shp)
    sv = np.zeros((ns,) + shp)
    node = np.clip(np.arange(-g, n_cells + g), 0, n_cells - 1)
    for side, (T, p, u, y) in enumerate((left, right)):
        y = [float(v) for v in y]
        R = 0.0
        for l in range(ns):
            R = R + y[l] * gas[l]
        rho = p / (R * T)
        cols = (node < n_left) if side == 0 else (node >= n_left)
        bv[0][..., cols] = rho
        bv[1][..., cols] = u
        bv[4][..., cols] = p
        bv[5][..., cols] = T
        for l in range(ns):
            sv[l][..., cols] = y[l]
    kinds = {"outflow": 6, "symmetry": 3, "wall": 3}
    faces = face_list(n_cells, ny, 1, {"i-": kinds[bc[0]], "i+": kinds[bc[1]], "j-": 6, "j+": 6})
    run = dict(mixture=1, n_spec=ns, n_var=5 + ns, n_scalar=ns, n_scalar_transported=ns, inviscid_scheme=scheme,
               viscous_order=0, reaction=reaction, gradp_in_diffusion=0, fixed_time_step=0, dt=0.0, cfl=cfl,
               prandtl=0.72, rho_ref=1.0, v_ref=400.0)
    x = np.zeros((3,) + shp)
    x[0] = np.arange(-g, n_cells + g) * dx
    return Case(n_cells, ny, 1, g, metric, jac, bv, sv, faces, run, x)


# ---------------------------------------------------------------------------
# 3-D periodic curvilinear box with a seeded H2/air field
# ---------------------------------------------------------------------------
def _fourier_field(rng, axes, L, kmax=8, nmodes=24):
    """Seeded sum of cosine modes with integer wavenumbers |k| <= kmax
    (periodic on the box) on the lattice given by 1-D axes (x, y, z); the
    result has shape (nz, ny, nx)."""
    x, y, z = axes
    f = np.zeros((z.size, y.size, x.size))
    for _ in range(nmodes):
        while True:
            k = rng.integers(-kmax, kmax + 1, size=3)
            if 0 < np.sqrt((k * k).sum()) <= kmax:
                break
        amp = rng.uniform(0.2, 1.0) / np.sqrt((k * k).sum())
        ph = rng.uniform(0, 2 * np.pi)
        ex = np.exp(1j * (2 * np.pi * k[0] * x / L[0] + ph))
        ey = np.exp(1j * 2 * np.pi * k[1] * y / L[1])
        ez = np.exp(1j * 2 * np.pi * k[2] * z / L[2])
        f += amp * (ez[:, None, None] * (ey[:, None] * ex[None, :])[None, :, :]).real
    return f


def _box_grid(mx, my, mz, g, rng, lengths, amp):
    """Periodic box lattice with a seeded sinusoidal node displacement of amp
    spacings; returns coordinates, metrics, the undisplaced axes and the
    interior slice."""
    Lx, Ly, Lz = lengths
    dx, dy, dz = Lx / (mx - 1), Ly / (my - 1), Lz / max(mz - 1, 1)
    ph = rng.uniform(0, 2 * np.pi, size=6)

    def coord(I, J, K):
        x0, y0, z0 = I * dx, J * dy, K * dz
        X = x0 + amp * dx * np.sin(2 * np.pi * y0 / Ly + ph[0]) * np.sin(2 * np.pi * z0 / Lz + ph[1])
        Y = y0 + amp * dy * np.sin(2 * np.pi * z0 / Lz + ph[2]) * np.sin(2 * np.pi * x0 / Lx + ph[3])
        Z = z0 + amp * dz * np.sin(2 * np.pi * x0 / Lx + ph[4]) * np.sin(2 * np.pi * y0 / Ly + ph[5])
        return X, Y, Z

    X, Y, Z = _grid_nodes(mx, my, mz, g, coord)
    metric, jac = metrics(X, Y, Z)
    # field values are functions of the undisplaced lattice position so that
    # coincident periodic nodes carry identical values
    pos = (np.arange(-g, mx + g) * dx, np.arange(-g, my + g) * dy, np.arange(-g, mz + g) * dz)
    interior = (slice(g, g + mz), slice(g, g + my), slice(g, g + mx))
    return X, Y, Z, metric, jac, pos, interior


def reacting_box_case(mx: int, my: int, mz: int, seed: int = 2024, tables: _abi.ChemTables | None = None,
                      lengths=(2 * np.pi * 1e-3, np.pi * 1e-3, np.pi * 1e-3), amp: float = 0.1,
                      T_range=(300.0, 1500.0), u_rms: float = 50.0) -> Case:
    """9-species H2/O2/N2 case on a sinusoidally perturbed periodic box.

    Nodes 0 and m-1 coincide through periodicity (the reference's periodic
    treatment averages them).  Node displacement 0.1 spacing, seeded phases;
    T in [300, 1500] K and p = 1 atm +-5% from seeded Fourier modes (k <= 8),
    mixture fraction Z in [0, 1] -> Y_H2 = Z, Y_O2 = 0.233 (1-Z),
    Y_N2 = 0.767 (1-Z); velocity modes scaled to an rms of 50 m/s.
    """
    g = 4
    rng = np.random.default_rng(seed)
    X, Y, Z, metric, jac, pos, interior = _box_grid(mx, my, mz, g, rng, lengths, amp)

    def norm01(f):
        fi = f[interior]
        return (f - fi.min()) / (fi.max() - fi.min())

    fT = norm01(_fourier_field(rng, pos, lengths))
    fp = norm01(_fourier_field(rng, pos, lengths))
    fZ = np.clip(norm01(_fourier_field(rng, pos, lengths)), 0.0, 1.0)
    vel = []
    for _ in range(3):
        f = _fourier_field(rng, pos, lengths)
        f = f - f[interior].mean()
        vel.append(f)
    rms = np.sqrt(np.mean(sum(v[interior] ** 2 for v in vel)))
    vel = [v * (u_rms / rms) for v in vel]

    shp = jac.shape
    T = T_range[0] + (T_range[1] - T_range[0]) * fT
    p = P_ATM * (1.0 + 0.05 * (2.0 * fp - 1.0))
    # species order of the Li-2004 mechanism: H2 H O2 O OH HO2 H2O2 H2O N2
    ns = 9
    sv = np.zeros((ns,) + shp)
    sv[0] = fZ
    sv[2] = 0.233 * (1.0 - fZ)
    sv[8] = 0.767 * (1.0 - fZ)
    if tables is None:
        mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
    else:
        mw = np.array(tables.mw[:ns])
    imw = (sv / mw[:, None, None, None]).sum(axis=0)
    rho = p / (R_U * imw * T)
    bv = np.zeros((6,) + shp)
    bv[0] = rho
    bv[1], bv[2], bv[3] = vel
    bv[4] = p
    bv[5] = T
    run = dict(mixture=1, n_spec=ns, n_var=5 + ns, n_scalar=ns, n_scalar_transported=ns, inviscid_scheme=72,
               viscous_order=8, reaction=1, gradp_in_diffusion=0, fixed_time_step=1, dt=0.0, cfl=0.5,
               prandtl=0.72, rho_ref=float(rho[interior].mean()), v_ref=u_rms)
    faces = face_list(mx, my, mz, {k: 8 for k in ("i-", "i+", "j-", "j+", "k-", "k+")})
    return Case(mx, my, mz, g, metric, jac, bv, sv, faces, run,
                np.stack([X[1:-1, 1:-1, 1:-1], Y[1:-1, 1:-1, 1:-1], Z[1:-1, 1:-1, 1:-1]]))


# ---------------------------------------------------------------------------
# 3-D periodic curvilinear box with a seeded, partly dissociated AIR-5 field
# ---------------------------------------------------------------------------
AIR5_MECH = DATA / "air5_park"


def air5_box_case(mx: int, my: int, mz: int, seed: int = 2025, tables: _abi.ChemTables | None = None,
                  two_temperature: bool = False, lengths=(2e-2, 1e-2, 1e-2), amp: float = 0.1,
                  T_range=(300.0, 9000.0), p0: float = 1.0e4, u_rms: float = 300.0, ngg: int = 3) -> Case:
    """5-species air (N2, O2, NO, N, O) on the perturbed periodic box of
    reacting_box_case.  Seeded Fourier fields give T in T_range (crossing the
    NASA-9 range cuts at 1000 K and 6000 K), p = p0 +-5%, dissociated
    fractions dN in [0, 0.4] and dO in [0, 0.8] (scaled by the temperature
    field) of the 0.767 / 0.233 N2 / O2 air, up to 4% NO, and velocity modes of rms u_rms.  With two_temperature,
    Tve ranges over [0.5 T, 1.5 T] and the E_ve scalar (n_scalar = 6) is left
    zero for the initialisation to derive from Tve.  WENO5, 2nd-order viscous
    terms, chemistry, fixed time step; ngg ghost layers (4 for WENO7)."""
    g = ngg
    rng = np.random.default_rng(seed)
    X, Y, Z, metric, jac, pos, interior = _box_grid(mx, my, mz, g, rng, lengths, amp)

    def norm01(f):
        fi = f[interior]
        return np.clip((f - fi.min()) / (fi.max() - fi.min()), 0.0, 1.0)

    fT = norm01(_fourier_field(rng, pos, lengths))
    fp = norm01(_fourier_field(rng, pos, lengths))
    fN = norm01(_fourier_field(rng, pos, lengths))
    fO = norm01(_fourier_field(rng, pos, lengths))
    fNO = norm01(_fourier_field(rng, pos, lengths))
    fV = norm01(_fourier_field(rng, pos, lengths))
    vel = []
    for _ in range(3):
        f = _fourier_field(rng, pos, lengths)
        f = f - f[interior].mean()
        vel.append(f)
    rms = np.sqrt(np.mean(sum(v[interior] ** 2 for v in vel)))
    vel = [v * (u_rms / rms) for v in vel]

    shp = jac.shape
    T = T_range[0] + (T_range[1] - T_range[0]) * fT
    p = p0 * (1.0 + 0.05 * (2.0 * fp - 1.0))
    ns = 5
    # dissociation grows with temperature, so cold regions hold little atomic N and O
    dN, dO, yNO = 0.4 * fN * fT, 0.8 * fO * fT, 0.04 * fNO
    sv = np.zeros((ns + (1 if two_temperature else 0),) + shp)
    sv[0] = 0.767 * (1 - dN) * (1 - yNO)
    sv[1] = 0.233 * (1 - dO) * (1 - yNO)
    sv[2] = yNO
    sv[3] = 0.767 * dN * (1 - yNO)
    sv[4] = 0.233 * dO * (1 - yNO)
    mw = np.array(tables.mw[:ns]) if tables is not None else np.array([28.0134, 31.9988, 30.0061, 14.0067, 15.9994])
    imw = (sv[:ns] / mw[:, None, None, None]).sum(axis=0)
    rho = p / (R_U * imw * T)
    bv = np.zeros((6,) + shp)
    bv[0] = rho
    bv[1], bv[2], bv[3] = vel
    bv[4] = p
    bv[5] = T
    nsc = ns + (1 if two_temperature else 0)
    run = dict(mixture=1, n_spec=ns, n_var=5 + nsc, n_scalar=nsc, n_scalar_transported=nsc, inviscid_scheme=52,
               viscous_order=2, reaction=1, gradp_in_diffusion=0, fixed_time_step=1, dt=0.0, cfl=0.5,
               prandtl=0.72, rho_ref=float(rho[interior].mean()), v_ref=u_rms,
               two_temperature=1 if two_temperature else 0, vt_relaxation=1 if two_temperature else 0)
    faces = face_list(mx, my, mz, {k: 8 for k in ("i-", "i+", "j-", "j+", "k-", "k+")})
    c = Case(mx, my, mz, g, metric, jac, bv, sv, faces, run,
             np.stack([X[1:-1, 1:-1, 1:-1], Y[1:-1, 1:-1, 1:-1], Z[1:-1, 1:-1, 1:-1]]))
    if two_temperature:
        c.tve = T * (0.5 + fV)
    return c


# ---------------------------------------------------------------------------
# 2-D O-grid around a square body: Mach-10 two-temperature air
# ---------------------------------------------------------------------------
def square_ogrid_case(mx: int = 1024, my: int = 512, seed: int = 4, side: float = 0.01, radius: float = 0.05,
                      ratio: float = 1.05, mach: float = 10.0, tables: _abi.ChemTables | None = None) -> Case:
    """O-grid (mz = 1) between a square of side `side` and a circle of
    `radius`, both centred at the origin.  i runs clockwise around the body
    (nodes 0 and mx-1 coincide: a periodic face pair), j runs outward with a
    wall-normal stretching ratio `ratio` (first spacing from the radial
    distance at the square's mid-side); the angular origin of i is seeded.
    Boundaries: j- isothermal 300 K wall (label 2), j+ freestream inflow
    (label 5) where the circle faces the flow (x < 0) and outflow (label 6)
    elsewhere, i faces periodic (label 8).  The field starts at the
    freestream: Mach `mach`, T = Tve = 220 K, rho = 0.02 kg/m^3, N2/O2
    0.767/0.233 by mass; two-temperature 5-species air, WENO5, 2nd-order
    viscous terms, chemistry, V-T relaxation, limit_flow, fixed time step.
    The velocity starts at rest on the wall and reaches the freestream at a
    third of the radial index range (smoothstep in the stretched distance)."""
    rng = np.random.default_rng(seed)
    g = 3
    phase = rng.uniform(0.0, 1.0)
    half = 0.5 * side
    nj = my - 1
    fr = (ratio ** np.arange(my) - 1.0) / (ratio ** nj - 1.0)

    def square(sp):
        # clockwise perimeter from (half, half) by arc length fraction sp in [0, 1)
        u = 4.0 * (sp % 1.0)
        side_i = np.floor(u)
        t = u - side_i
        x = np.select([side_i == 0, side_i == 1, side_i == 2], [half + 0 * t, half - 2 * half * t, -half + 0 * t],
                      -half + 2 * half * t)
        y = np.select([side_i == 0, side_i == 1, side_i == 2], [half - 2 * half * t, -half + 0 * t, -half + 2 * half * t],
                      half + 0 * t)
        return x, y

    def coord(I, J, K):
        sp = I / (mx - 1) + phase
        xs, ys = square(sp)
        th = np.pi / 4 - 2 * np.pi * sp           # clockwise, the corner (half, half) at 45 degrees
        xc, yc = radius * np.cos(th), radius * np.sin(th)
        Jc = np.clip(J, 0, nj).astype(int)
        f = fr[Jc]
        f = np.where(J < 0, J * fr[1], f)
        f = np.where(J > nj, fr[nj] + (J - nj) * (fr[nj] - fr[nj - 1]), f)
        X = (1 - f) * xs + f * xc
        Y = (1 - f) * ys + f * yc
        return X, Y, K * 1e-3

    X, Y, Z = _grid_nodes(mx, my, 1, g, coord)
    metric, jac = metrics(X, Y, Z)
    shp = jac.shape
    ns = 5
    mw = np.array(tables.mw[:ns]) if tables is not None else np.array([28.0134, 31.9988, 30.0061, 14.0067, 15.9994])
    y_in = np.array([0.767, 0.233, 0.0, 0.0, 0.0])
    mw_in = 1.0 / (y_in / mw).sum()
    T_in, rho_in = 220.0, 0.02
    p_in = rho_in * R_U / mw_in * T_in
    u_in = mach * np.sqrt(1.4 * p_in / rho_in)
    bv = np.zeros((6,) + shp)
    # the velocity rises smoothly from rest at the wall to the freestream over the inner third
    jj = np.arange(-g, my + g)
    fj = np.clip(np.interp(jj, np.arange(my), fr) / fr[nj // 3], 0.0, 1.0)
    ramp = (fj * fj * (3 - 2 * fj))[None, :, None] * np.ones(shp)
    bv[0], bv[1], bv[4], bv[5] = rho_in, u_in * ramp, p_in, T_in
    sv = np.zeros((ns + 1,) + shp)
    sv[:ns] = y_in[:, None, None, None]
    # j+ face: inflow where the circle faces the flow, outflow elsewhere (contiguous ranges)
    xs_out = radius * np.cos(np.pi / 4 - 2 * np.pi * (np.arange(mx) / (mx - 1) + phase))
    label = np.where(xs_out < 0, 5, 6)
    faces = [(0, -1, 8, (0, 0, 0), (0, my - 1, 0)), (0, 1, 8, (mx - 1, 0, 0), (mx - 1, my - 1, 0)),
             (1, -1, 2, (0, 0, 0), (mx - 1, 0, 0))]
    start = 0
    for i in range(1, mx + 1):
        if i == mx or label[i] != label[start]:
            faces.append((1, 1, int(label[start]), (start, my - 1, 0), (i - 1, my - 1, 0)))
            start = i
    inflow = dict(label=5, kind=5, inflow_sub=0, density=rho_in, u=u_in, v=0.0, w=0.0, pressure=p_in,
                  temperature=T_in, mw=mw_in, sv=list(y_in) + [0.0], temperature_ve=T_in)
    conditions = [inflow, dict(label=2, kind=2, thermal_type=0, wall_temperature=300.0),
                  dict(label=6, kind=6), dict(label=8, kind=8)]
    run = dict(mixture=1, n_spec=ns, n_var=5 + ns + 1, n_scalar=ns + 1, n_scalar_transported=ns + 1,
               inviscid_scheme=52, viscous_order=2, reaction=1, gradp_in_diffusion=0, fixed_time_step=1, dt=0.0,
               cfl=0.5, prandtl=0.72, rho_ref=rho_in, v_ref=u_in, p_ref=p_in, limit_flow=1, two_temperature=1,
               vt_relaxation=1)
    c = Case(mx, my, 1, g, metric, jac, np.ascontiguousarray(bv), np.ascontiguousarray(sv), faces, run,
             np.stack([X[1:-1, 1:-1, 1:-1], Y[1:-1, 1:-1, 1:-1], Z[1:-1, 1:-1, 1:-1]]), conditions=conditions)
    c.tve = bv[5].copy()
    return c


# ---------------------------------------------------------------------------
# Multi-block case: a box of blocks with 1-to-1 interfaces, some blocks rotated
# ---------------------------------------------------------------------------
# proper rotations of the local axes (signed permutations with determinant +1)
ROTATIONS = [((0, 1, 2), (1, 1, 1)), ((1, 0, 2), (-1, 1, 1)), ((0, 2, 1), (1, -1, 1)), ((2, 1, 0), (1, 1, -1)),
             ((1, 2, 0), (1, 1, 1)), ((0, 1, 2), (-1, -1, 1))]


def _inner_numbers(pa, sa, ma, a_n, side_a, pb, sb, mb, b_n, side_b):
    """Source / target lines of one interface in the inner%4d.txt notation.

    pa/sa: local->global axis map and signs of the source block, ma its local
    sizes, a_n/side_a the local normal axis and side (-1 low, +1 high) of the
    source face; the same for the target block."""
    tang = [a for a in range(3) if a != a_n]
    src = [0] * 6
    tar = [0] * 6
    src[2 * a_n] = src[2 * a_n + 1] = 1 if side_a < 0 else ma[a_n]
    tar[2 * b_n] = tar[2 * b_n + 1] = 1 if side_b < 0 else mb[b_n]
    for n, a in enumerate(tang):
        b = pb.index(pa[a])                      # target axis along the same global axis
        same = sa[a] == sb[b]
        mark = -1 if n == 0 else 1               # the first tangential axis carries the negative mark
        src[2 * a], src[2 * a + 1] = mark * 1, mark * ma[a]
        if same:
            tar[2 * b], tar[2 * b + 1] = mark * 1, mark * mb[b]
        else:
            tar[2 * b], tar[2 * b + 1] = mark * mb[b], mark * 1
    return src, tar


def multiblock_case(split=(3, 2, 2), widths=None, seed: int = 7, physics: str = "air", rotate: bool = True,
                    scale: float = 1.0, wiggle: float = 0.1, ngg: int | None = None) -> list:
    """Blocks tiling a box; neighbouring blocks share their interface nodes.

    widths[a][b] is the node count of block column b along global axis a.
    physics "air": Air perfect gas, characteristic WENO5, inviscid, outflow
    outer faces, a pressure pulse and density blob with seeded velocity
    modes.  physics "h2": the 9-species field of reacting_box_case with
    WENO7, 8th-order viscous terms and chemistry.  With rotate, some blocks
    carry rotated local axes, so interfaces connect different local faces
    and running directions.  wiggle is the node displacement in spacings;
    ngg overrides the ghost layers (3 for air, 4 for h2).
    Returns the list of block Cases."""
    rng = np.random.default_rng(seed)
    nb = split
    if widths is None:
        widths = [[int(v) for v in rng.integers(10, 17, size=nb[a])] for a in range(3)]
    starts = [np.concatenate([[0], np.cumsum([w - 1 for w in widths[a][:-1]])]).astype(int) for a in range(3)]
    N = [int(starts[a][-1] + widths[a][-1]) for a in range(3)]
    air = physics == "air"
    g = ngg if ngg is not None else (3 if air else 4)
    L = (1.0 * scale, 0.8 * scale, 0.6 * scale) if air else (2 * np.pi * 1e-3, np.pi * 1e-3, np.pi * 1e-3)
    d = [L[a] / (N[a] - 1) for a in range(3)]
    ph = rng.uniform(0, 2 * np.pi, size=6)

    def coord(G):
        x0, y0, z0 = G[0] * d[0], G[1] * d[1], G[2] * d[2]
        X = x0 + wiggle * d[0] * np.sin(2 * np.pi * y0 / L[1] + ph[0]) * np.sin(2 * np.pi * z0 / L[2] + ph[1])
        Y = y0 + wiggle * d[1] * np.sin(2 * np.pi * z0 / L[2] + ph[2]) * np.sin(2 * np.pi * x0 / L[0] + ph[3])
        Z = z0 + wiggle * d[2] * np.sin(2 * np.pi * x0 / L[0] + ph[4]) * np.sin(2 * np.pi * y0 / L[1] + ph[5])
        return X, Y, Z

    # global fields, evaluated at lattice positions
    if air:
        c0 = [0.45 * L[0], 0.5 * L[1], 0.55 * L[2]]
        kv = [rng.uniform(0.5, 2.0, size=3) for _ in range(3)]
        pv = [rng.uniform(0, 2 * np.pi, size=3) for _ in range(3)]

        def fields(P):
            r2 = sum((P[a] - c0[a]) ** 2 for a in range(3)) / (0.15 * L[0]) ** 2
            rho = 1.0 + 0.5 * np.exp(-r2)
            p = 1.0 + 2.0 * np.exp(-2.0 * r2)
            vel = [0.1 * np.sin(2 * np.pi * kv[c][0] * P[0] / L[0] + pv[c][0]) *
                   np.cos(2 * np.pi * kv[c][1] * P[1] / L[1] + pv[c][1]) *
                   np.sin(2 * np.pi * kv[c][2] * P[2] / L[2] + pv[c][2]) for c in range(3)]
            T = p * MW_AIR / (rho * R_U)
            return np.stack([rho, vel[0], vel[1], vel[2], p, T]), np.zeros((0,) + rho.shape)
        run = dict(mixture=0, n_spec=0, n_var=5, n_scalar=0, n_scalar_transported=0, inviscid_scheme=52,
                   viscous_order=0, reaction=0, gradp_in_diffusion=0, fixed_time_step=0, dt=0.0, cfl=0.5,
                   prandtl=0.72, rho_ref=1.0, v_ref=float(np.sqrt(1.4)))
    else:
        # coarse global lattice of Fourier modes, sampled per block
        lat = (np.arange(-g, N[0] + g) * d[0], np.arange(-g, N[1] + g) * d[1], np.arange(-g, N[2] + g) * d[2])
        inner_sl = (slice(g, g + N[2]), slice(g, g + N[1]), slice(g, g + N[0]))

        def norm01(f):
            fi = f[inner_sl]
            return (f - fi.min()) / (fi.max() - fi.min())
        fT = norm01(_fourier_field(rng, lat, L))
        fp = norm01(_fourier_field(rng, lat, L))
        fZ = np.clip(norm01(_fourier_field(rng, lat, L)), 0.0, 1.0)
        vel = []
        for _ in range(3):
            f = _fourier_field(rng, lat, L)
            vel.append(f - f[inner_sl].mean())
        rms = np.sqrt(np.mean(sum(v[inner_sl] ** 2 for v in vel)))
        vel = [v * (50.0 / rms) for v in vel]
        mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])

        def fields_at(G):
            ix = tuple(np.clip(G[2 - a] + g, 0, N[2 - a] + 2 * g - 1) for a in range(3))
            T = 300.0 + 1200.0 * fT[ix]
            p = P_ATM * (1.0 + 0.05 * (2.0 * fp[ix] - 1.0))
            Zf = fZ[ix]
            sv = np.zeros((9,) + T.shape)
            sv[0], sv[2], sv[8] = Zf, 0.233 * (1.0 - Zf), 0.767 * (1.0 - Zf)
            rho = p / (R_U * (sv / mw[:, None, None, None]).sum(axis=0) * T)
            return np.stack([rho, vel[0][ix], vel[1][ix], vel[2][ix], p, T]), sv
        run = dict(mixture=1, n_spec=9, n_var=14, n_scalar=9, n_scalar_transported=9, inviscid_scheme=72,
                   viscous_order=8, reaction=1, gradp_in_diffusion=0, fixed_time_step=1, dt=0.0, cfl=0.5,
                   prandtl=0.72, rho_ref=0.5, v_ref=50.0)

    bids = [(bx, by, bz) for bz in range(nb[2]) for by in range(nb[1]) for bx in range(nb[0])]
    orient = []
    for n, _ in enumerate(bids):
        orient.append(ROTATIONS[(n * 5 + seed) % len(ROTATIONS)] if rotate and n % 2 == 1 else ROTATIONS[0])
    blocks = []
    for n, bc in enumerate(bids):
        perm, sgn = orient[n]
        gmin = [int(starts[a][bc[a]]) for a in range(3)]
        gmax = [gmin[a] + widths[a][bc[a]] - 1 for a in range(3)]
        m = [widths[perm[a]][bc[perm[a]]] for a in range(3)]

        def to_global(li, lj, lk, perm=perm, sgn=sgn, gmin=gmin, gmax=gmax):
            loc = (li, lj, lk)
            G = [None] * 3
            for a in range(3):
                G[perm[a]] = gmin[perm[a]] + loc[a] if sgn[a] > 0 else gmax[perm[a]] - loc[a]
            return G
        kk, jj, ii = np.meshgrid(np.arange(-g - 1, m[2] + g + 1), np.arange(-g - 1, m[1] + g + 1),
                                 np.arange(-g - 1, m[0] + g + 1), indexing="ij")
        G = to_global(ii, jj, kk)
        X, Y, Z = coord([G[0].astype(float), G[1].astype(float), G[2].astype(float)])
        metric, jac = metrics(X, Y, Z)
        Gi = [a[1:-1, 1:-1, 1:-1] for a in G]
        if air:
            bv, sv = fields([Gi[a] * d[a] for a in range(3)])
        else:
            bv, sv = fields_at(Gi)
        faces = []
        inner = []
        for a_n in range(3):
            for side in (-1, 1):
                ga = perm[a_n]
                # global position of this face and the neighbour block across it
                at_max = (side > 0) == (sgn[a_n] > 0)
                nbc = list(bc)
                nbc[ga] += 1 if at_max else -1
                start = [0, 0, 0]
                end = [m[0] - 1, m[1] - 1, m[2] - 1]
                start[a_n] = end[a_n] = 0 if side < 0 else m[a_n] - 1
                if not (0 <= nbc[ga] < nb[ga]):
                    faces.append((a_n, side, 6, start, end))
                    continue
                t = bids.index(tuple(nbc))
                pb, sb = orient[t]
                mb = [widths[pb[a]][nbc[pb[a]]] for a in range(3)]
                b_n = pb.index(ga)
                # the neighbour touches this face with its face at the opposite global side
                nb_at_max = not at_max
                side_b = 1 if nb_at_max == (sb[b_n] > 0) else -1
                src, tar = _inner_numbers(perm, sgn, m, a_n, side, pb, sb, mb, b_n, side_b)
                inner.append((src, tar, t + 1))
        c = Case(m[0], m[1], m[2], g, metric, jac, np.ascontiguousarray(bv), np.ascontiguousarray(sv), faces,
                 dict(run), np.stack([X[1:-1, 1:-1, 1:-1], Y[1:-1, 1:-1, 1:-1], Z[1:-1, 1:-1, 1:-1]]), inner,
                 (perm, sgn, gmin, gmax))
        blocks.append(c)
    return blocks


# ---------------------------------------------------------------------------
# Channel with inflow, outflow, walls and symmetry planes
# ---------------------------------------------------------------------------
def channel_case(mx: int = 24, my: int = 16, mz: int = 12, physics: str = "air", mach: float = 2.0,
                 subsonic: bool = False, seed: int = 3, wiggle: float = 0.1) -> Case:
    """Curvilinear channel: i- inflow (label 5), i+ outflow (6), j- isothermal
    wall (label 2, 350 K), j+ adiabatic wall (label 21), k- and k+ symmetry
    planes (label 3).  The inflow carries the free stream (Mach `mach`, or
    0.3 with `subsonic` and the characteristic inflow treatment); the
    interior starts from the free stream with seeded perturbations.
    physics "air": perfect gas, WENO5, inviscid; "h2": 9 species, WENO7,
    8th-order viscous terms and chemistry; "air5": 5-species air at 2000 K
    with the two-temperature model (NASA-9, Park chemistry, V-T relaxation),
    WENO5, 2nd-order viscous terms, the stream's Tve 0.8 T and an interior
    Tve of 0.7-1.1 T."""
    rng = np.random.default_rng(seed)
    air = physics == "air"
    air5 = physics == "air5"
    g = 3 if (air or air5) else 4
    L = (2.0, 1.0, 0.75) if air else ((4e-2, 2e-2, 1.5e-2) if air5 else (4e-3, 2e-3, 1.5e-3))
    d = (L[0] / (mx - 1), L[1] / (my - 1), L[2] / max(mz - 1, 1))
    ph = rng.uniform(0, 2 * np.pi, size=3)

    def coord(I, J, K):
        x0, y0, z0 = I * d[0], J * d[1], K * d[2]
        X = x0 + wiggle * d[0] * np.sin(2 * np.pi * y0 / L[1] + ph[0]) * np.sin(np.pi * z0 / L[2])
        Y = y0 + wiggle * d[1] * np.sin(2 * np.pi * x0 / L[0] + ph[1]) * np.sin(np.pi * z0 / L[2] + ph[2])
        return X, Y, z0

    X, Y, Z = _grid_nodes(mx, my, mz, g, coord)
    metric, jac = metrics(X, Y, Z)
    shp = jac.shape
    xc, yc, zc = X[1:-1, 1:-1, 1:-1], Y[1:-1, 1:-1, 1:-1], Z[1:-1, 1:-1, 1:-1]
    bump = np.exp(-(((xc - 0.5 * L[0]) / (0.2 * L[0])) ** 2 + ((yc - 0.5 * L[1]) / (0.2 * L[1])) ** 2))
    M = 0.3 if subsonic else mach
    if air:
        rho0, p0 = 1.0, 1.0 / 1.4
        T0 = p0 * MW_AIR / (rho0 * R_U)
        c0 = np.sqrt(1.4 * p0 / rho0)
        sv_in = []
        mw_in = MW_AIR
        sv = np.zeros((0,) + shp)
    elif air5:
        T0, p0 = 2000.0, 1.0e4
        mw = np.array([28.0134, 31.9988, 30.0061, 14.0067, 15.9994])
        y_in = np.array([0.767, 0.233, 0.0, 0.0, 0.0])
        mw_in = 1.0 / (y_in / mw).sum()
        rho0 = p0 * mw_in / (R_U * T0)
        c0 = np.sqrt(1.33 * p0 / rho0)
        sv_in = list(y_in) + [0.0]
        sv = np.zeros((6,) + shp)
        sv[:5] = y_in[:, None, None, None]
        sv[3] += 0.02 * bump
        sv[0] -= 0.02 * bump
    else:
        T0, p0 = 900.0, P_ATM
        mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
        y_in = np.zeros(9)
        y_in[0], y_in[2], y_in[8] = 0.02, 0.233 * 0.98, 0.767 * 0.98
        mw_in = 1.0 / (y_in / mw).sum()
        rho0 = p0 * mw_in / (R_U * T0)
        c0 = np.sqrt(1.35 * p0 / rho0)
        sv_in = list(y_in)
        sv = np.broadcast_to(y_in[:, None, None, None], (9,) + shp).copy()
        sv[0] += 0.02 * bump
        sv[8] -= 0.02 * bump
    U = M * c0
    bv = np.zeros((6,) + shp)
    bv[0] = rho0 * (1 + 0.2 * bump)
    bv[1] = U * (1 + 0.05 * np.sin(3 * xc / L[0] + ph[0]))
    bv[2] = 0.05 * U * np.sin(2 * np.pi * xc / L[0]) * bump
    bv[3] = 0.02 * U * np.cos(np.pi * zc / L[2]) * bump
    bv[4] = p0 * (1 + 0.3 * bump)
    if air:
        bv[5] = bv[4] * MW_AIR / (bv[0] * R_U)
    else:
        ns_ = len(mw)
        mwm = 1.0 / (sv[:ns_] / mw[:, None, None, None]).sum(axis=0)
        bv[5] = bv[4] * mwm / (bv[0] * R_U)
    inflow = dict(label=5, kind=5, inflow_sub=int(subsonic), density=rho0, u=U, v=0.0, w=0.0, pressure=p0,
                  temperature=T0, mw=mw_in, sv=sv_in)
    if air5:
        inflow["temperature_ve"] = 0.8 * T0
    conditions = [inflow, dict(label=2, kind=2, thermal_type=0,
                               wall_temperature=1.2 * T0 if air else (1500.0 if air5 else 350.0)),
                  dict(label=21, kind=2, thermal_type=1), dict(label=3, kind=3), dict(label=6, kind=6)]
    faces = face_list(mx, my, mz, {"i-": 5, "i+": 6, "j-": 2, "j+": 21, "k-": 3, "k+": 3})
    if air:
        run = dict(mixture=0, n_spec=0, n_var=5, n_scalar=0, n_scalar_transported=0, inviscid_scheme=52,
                   viscous_order=0, reaction=0, gradp_in_diffusion=0, fixed_time_step=0, dt=0.0, cfl=0.5,
                   prandtl=0.72, rho_ref=rho0, v_ref=U)
    elif air5:
        run = dict(mixture=1, n_spec=5, n_var=11, n_scalar=6, n_scalar_transported=6, inviscid_scheme=52,
                   viscous_order=2, reaction=1, gradp_in_diffusion=0, fixed_time_step=1, dt=0.0, cfl=0.5,
                   prandtl=0.72, rho_ref=rho0, v_ref=U, two_temperature=1, vt_relaxation=1)
    else:
        run = dict(mixture=1, n_spec=9, n_var=14, n_scalar=9, n_scalar_transported=9, inviscid_scheme=72,
                   viscous_order=8, reaction=1, gradp_in_diffusion=0, fixed_time_step=1, dt=0.0, cfl=0.5,
                   prandtl=0.72, rho_ref=rho0, v_ref=U)
    c = Case(mx, my, mz, g, metric, jac, np.ascontiguousarray(bv), np.ascontiguousarray(sv), faces, run,
             np.stack([xc, yc, zc]), conditions=conditions)
    if air5:
        c.tve = bv[5] * (0.7 + 0.4 * bump)
    return c
