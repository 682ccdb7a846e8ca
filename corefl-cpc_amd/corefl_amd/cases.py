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
    Lx, Ly, Lz = lengths
    dx, dy, dz = Lx / (mx - 1), Ly / (my - 1), Lz / (mz - 1)
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
