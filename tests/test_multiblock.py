"""Multi-block cases: interface registration and the oracle's exchange.

Registration (InnerFace::register_boundary, src/Mesh.cpp:39-172) is pinned
two ways: on generated 3-D block boxes with rotated local axes, every face
and ghost node it maps must land on the same global lattice node; on the
reference's own 24-block inner-face file (the square-cylinder example) the
maps must stay inside the target block, keep the normal distance, and be
mutual between the two sides.  The exchange (inner_communication,
src/DataCommunication.cu:37-124, and inner_exchange task 1, 489-541) is pinned
by running the same global field as one block and as twelve blocks (some
rotated): interfaces must be invisible.
"""
from pathlib import Path

import numpy as np
import pytest

from corefl_amd.cases import multiblock_case
from oracle import Oracle, register_inner

GOLDEN = Path(__file__).resolve().parent / "golden" / "multiblock"


def face_map(f):
    """All loop positions of a registered face -> (source idx, target idx)."""
    out = []
    for n2 in range(f.n_point[2]):
        for n1 in range(f.n_point[1]):
            for n0 in range(f.n_point[0]):
                n = (n0, n1, n2)
                d = [f.loop_dir[a] * n[a] for a in range(3)]
                idx = tuple(f.range_start[a] + d[a] for a in range(3))
                tar = tuple(f.target_start[a] + f.target_loop_dir[a] * d[f.src_tar[a]] for a in range(3))
                out.append((idx, tar))
    return out


def to_global(case, idx):
    perm, sgn, gmin, gmax = case.lattice
    G = [0, 0, 0]
    for a in range(3):
        G[perm[a]] = gmin[perm[a]] + idx[a] if sgn[a] > 0 else gmax[perm[a]] - idx[a]
    return tuple(G)


@pytest.mark.parametrize("seed", [3, 7, 11])
def test_registration_maps_coincident_nodes_3d(seed):
    blocks = multiblock_case(split=(3, 2, 2), seed=seed, physics="air", rotate=True)
    n_faces = 0
    for b, case in enumerate(blocks):
        for src, tar, tblk in case.inner:
            f = register_inner(src, tar, tblk, 3, case.ngg)
            target = blocks[f.target_block]
            m = (case.mx, case.my, case.mz)
            pairs = face_map(f)
            # the loop covers the face node and ngg ghost layers, tangentially the whole face
            assert len(pairs) == (case.ngg + 1) * np.prod([m[a] for a in range(3) if a != f.face])
            mt = (target.mx, target.my, target.mz)
            for idx, it in pairs:
                assert to_global(case, idx) == to_global(target, it)
                assert all(0 <= it[a] < mt[a] for a in range(3))
            n_faces += 1
    assert n_faces == 2 * (2 * 2 * 2 + 3 * 2 * 1 + 3 * 2 * 1)  # interfaces counted from both sides


def read_inner_file(path):
    """boundary_condition/inner%4d.txt as read by Mesh::read_inner_interface (src/Mesh.cpp:927-958)."""
    lines = path.read_text().splitlines()
    counts = [int(v) for v in lines[1].split()]
    faces, pos = [], 3
    for blk, n in enumerate(counts):
        for _ in range(n):
            a = [int(v) for v in lines[pos].split()]
            b = [int(v) for v in lines[pos + 1].split()]
            faces.append((blk, a[:6], b[:6], b[6]))
            pos += 2
    return counts, faces


def test_registration_reference_square_cylinder_2d():
    counts, faces = read_inner_file(GOLDEN / "square_cylinder_inner.txt")
    assert len(counts) == 24 and sum(counts) == len(faces) == 72
    # block extents from the face ranges (each block touches its max faces)
    ext = np.zeros((24, 2), int)
    for blk, a, b, tb in faces:
        ext[blk] = np.maximum(ext[blk], [max(abs(a[0]), abs(a[1])), max(abs(a[2]), abs(a[3]))])
        ext[tb - 1] = np.maximum(ext[tb - 1], [max(abs(b[0]), abs(b[1])), max(abs(b[2]), abs(b[3]))])
    ngg = 3
    reg = [(blk, register_inner(a, b, tb, 2, ngg)) for blk, a, b, tb in faces]
    coincident = {}
    for blk, f in reg:
        m = (ext[blk][0], ext[blk][1], 1)
        tgt = f.target_block
        mt = (ext[tgt][0], ext[tgt][1], 1)
        face_src = f.range_start[f.face] if f.direction > 0 else f.range_end[f.face]
        face_tar = 0 if f.target_direction < 0 else mt[f.target_face] - 1
        assert f.n_point[f.face] == ngg + 1 and f.n_point[2] == 1
        assert f.n_point[1 - f.face] == m[1 - f.face]
        for idx, it in face_map(f):
            assert all(0 <= it[a] < mt[a] for a in range(2)), (blk, idx, it)
            assert it[2] == 0
            # the normal distance from the face is kept: ghost layer d lands d nodes inside the target
            assert abs(idx[f.face] - face_src) == abs(it[f.target_face] - face_tar)
            if idx[f.face] == face_src:
                coincident.setdefault((blk, idx[:2]), set()).add((tgt, it[:2]))
    # the two sides of every interface map their face nodes onto each other
    # (corner nodes lie on two interfaces)
    for (blk, node), targets in coincident.items():
        for tgt, tnode in targets:
            assert (blk, node) in coincident[(tgt, tnode)]


def _global_field(oracle_blocks, cases, key):
    """Scatter each block's interior values of bv/cv planes onto the global lattice."""
    N = [max(max(to_global(c, (i, j, k))[a] for i in (0, c.mx - 1) for j in (0, c.my - 1) for k in (0, c.mz - 1))
             for c in cases) + 1 for a in range(3)]
    nvar = oracle_blocks[0].a[key].shape[0]
    out = np.full((nvar, N[2], N[1], N[0]), np.nan)
    for ob, c in zip(oracle_blocks, cases):
        g = c.ngg
        arr = ob.a[key][:, g:g + c.mz, g:g + c.my, g:g + c.mx]
        kk, jj, ii = np.meshgrid(np.arange(c.mz), np.arange(c.my), np.arange(c.mx), indexing="ij")
        perm, sgn, gmin, gmax = c.lattice
        loc = (ii, jj, kk)
        G = [None] * 3
        for a in range(3):
            G[perm[a]] = gmin[perm[a]] + loc[a] if sgn[a] > 0 else gmax[perm[a]] - loc[a]
        out[:, G[2], G[1], G[0]] = arr
    return out


@pytest.mark.parametrize("physics,steps,rotate,wiggle,tol", [
    ("air", 3, False, 0.1, 1e-12),
    ("air", 3, True, 0.0, 1e-12),
    ("h2", 1, False, 0.1, 1e-9),
])
def test_interfaces_are_invisible(physics, steps, rotate, wiggle, tol, h2_oracle_tables):
    """Twelve blocks against one block of the same field (conservative variables).

    Rotated blocks are compared on a Cartesian grid only: on a curvilinear
    grid the reference's WENO face metric (averaged along local i in every
    sweep, src/WENO.cu:1615-1617, 2280-2282) depends on the local axis order,
    and its viscous-flux exchange copies xi/eta/zeta fluxes component by
    component (src/DataCommunication.cu:507-527), so a rotated block is a
    different discretisation there.  For the mixture, interface nodes get cv
    back from (rho, u, p, T) after the exchange, and T carries the Newton
    tolerance of the update (relative 1e-5 on the last correction), which
    bounds the agreement at about 1e-11 per stage."""
    widths = [[9, 10, 9], [9, 10], [10, 9]]
    multi = multiblock_case(split=(3, 2, 2), widths=widths, seed=5, physics=physics, rotate=rotate, wiggle=wiggle)
    N = [sum(w) - (len(w) - 1) for w in widths]
    single = multiblock_case(split=(1, 1, 1), widths=[[N[0]], [N[1]], [N[2]]], seed=5, physics=physics,
                             rotate=False, wiggle=wiggle)
    tables = h2_oracle_tables if physics == "h2" else None
    om, os_ = Oracle(multi, tables), Oracle(single, tables)
    om.initialize()
    os_.initialize()
    dt = 0.5 * os_.cfl_dt()
    for _ in range(steps):
        om.step(dt, True)
        os_.step(dt, True)
    # cv momentum components are global-frame in every block: compare directly
    cm = _global_field(om.blocks, multi, "cv")
    cs = _global_field(os_.blocks, single, "cv")
    assert not np.isnan(cm).any()
    mom = np.abs(cs[1:4]).max()
    for l in range(cs.shape[0]):
        scale = mom if 1 <= l <= 3 else np.abs(cs[l]).max()
        err = np.abs(cm[l] - cs[l]).max() / scale
        assert err <= tol, (l, err)
