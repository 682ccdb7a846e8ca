"""Solution files (corefl_plt_write / corefl_plt_read): the reference's
output/flowfield.plt (Tecplot binary V112 written by FieldIO<>,
src/FieldIO.h:96-420 and 709-800) and its restart reader (read_flowfield,
src/Initialize.cu:167-420).

The file is parsed here from Tecplot's published V112 layout (header: magic,
byte order, file type, title, variable names, zone records, end-of-header
marker 357.0; data section per zone: marker 299.0, one format word per
variable, passive / shared / connectivity words, min-max pairs, then the
variables, i fastest), independently of the library's reader.  Checks:
every field bit for bit against the solver's own planes, the min / max pairs
against numpy, a write -> read round trip bit for bit, zones matched to
blocks by node counts when the block order differs, loud errors where the
reference would silently interpolate or lacks a variable, and a restarted
run against the oracle restarted the same way (cv rebuilt from bv) at 1e-10
normwise after two further steps."""
import struct

import numpy as np
import pytest

from corefl_amd import CoreflError, Solver, _abi, load_tables, reacting_box_case
from oracle import Oracle, h2_tables

pytestmark = pytest.mark.gpu

SPECIES = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]


@pytest.fixture(scope="module")
def tables():
    return load_tables()


class V112:
    """Reads a Tecplot V112 binary file of ordered, all-double zones."""

    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.o)
        self.o += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def string(self):
        s = ""
        while True:
            c = self.take("i")
            if c == 0:
                return s
            s += chr(c)

    def parse(self):
        assert self.b[:8] == b"#!TDV112"
        self.o = 8
        assert self.take("i") == 1  # byte order
        assert self.take("i") == 0  # full file
        title = self.string()
        n_var = self.take("i")
        names = [self.string() for _ in range(n_var)]
        zones = []
        marker = self.take("f")
        while marker == 299.0:
            name = self.string()
            parent, strand, time, colour, ztype, loc, raw, misc = self.take("iidiiiii")
            dims = self.take("iii")
            assert self.take("i") == 0
            zones.append(dict(name=name, parent=parent, strand=strand, time=time, colour=colour, type=ztype,
                              loc=loc, raw=raw, misc=misc, dims=dims))
            marker = self.take("f")
        assert marker == 357.0
        for z in zones:
            assert self.take("f") == 299.0
            assert list(self.take(f"{n_var}i") if n_var > 1 else [self.take("i")]) == [2] * n_var
            assert self.take("iii") == (0, 0, -1)
            z["minmax"] = np.array(self.take(f"{2 * n_var}d")).reshape(n_var, 2)
            n = int(np.prod(z["dims"]))
            z["data"] = np.frombuffer(self.b, dtype="<f8", count=n * n_var, offset=self.o).reshape(
                n_var, z["dims"][2], z["dims"][1], z["dims"][0])
            self.o += n * n_var * 8
        assert self.o == len(self.b)
        return title, names, zones


def box(tables, mx, my, mz, seed):
    return reacting_box_case(mx, my, mz, seed=seed, tables=tables)


def stepped(tables, case, steps, dt):
    s = Solver(case.run, tables)
    s.add_block(case)
    s.initialize()
    for _ in range(steps):
        s.step(dt)
    return s


@pytest.mark.parametrize("ngg_out", [0, 2])
def test_file_layout_and_values(tables, tmp_path, ngg_out):
    case = box(tables, 12, 10, 8, seed=3)
    s = stepped(tables, case, 1, 2e-9)
    path = tmp_path / "flowfield.plt"
    print(f"plt layout ngg_out {ngg_out}", flush=True)
    s.write_plt(path, ngg_out=ngg_out, solution_time=1.25e-6, step_path=tmp_path / "step.txt", step=7)
    title, names, zones = V112(path).parse()
    assert title == "Solution file"
    assert names == ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", *SPECIES, "mach"]
    assert len(zones) == 1
    z = zones[0]
    assert z["name"] == "zone 0" and z["time"] == 1.25e-6
    assert (z["parent"], z["strand"], z["colour"], z["type"], z["loc"], z["raw"], z["misc"]) == (-1, -2, -1, 0, 0, 0, 0)
    assert z["dims"] == (12 + 2 * ngg_out, 10 + 2 * ngg_out, 8 + 2 * ngg_out)
    g = case.ngg - ngg_out
    sl = (slice(g, -g), slice(g, -g), slice(g, -g))
    want = np.concatenate([case.x[(slice(None),) + sl], s.download(0, _abi.F_BV)[(slice(None),) + sl],
                           s.download(0, _abi.F_SV)[(slice(None),) + sl],
                           s.download(0, _abi.F_MACH).reshape(1, *case.padded_shape)[(slice(None),) + sl]])
    assert np.array_equal(z["data"], want)
    assert np.array_equal(z["minmax"][:, 0], want.reshape(len(names), -1).min(axis=1))
    assert np.array_equal(z["minmax"][:, 1], want.reshape(len(names), -1).max(axis=1))
    assert (tmp_path / "step.txt").read_text() == "7"
    s.close()


def test_round_trip_is_bit_exact(tables, tmp_path):
    case = box(tables, 12, 10, 8, seed=4)
    a = stepped(tables, case, 2, 2e-9)
    path = tmp_path / "flowfield.plt"
    a.write_plt(path, solution_time=4e-9)
    b = Solver(case.run, tables)
    b.add_block(box(tables, 12, 10, 8, seed=5))  # another state, same grid
    ghosts_before = b.download(0, _abi.F_BV)
    print("plt round trip", flush=True)
    assert b.read_plt(path) == 4e-9
    g = case.ngg
    inner = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    for f in (_abi.F_BV, _abi.F_SV):
        assert np.array_equal(b.download(0, f)[inner], a.download(0, f)[inner])
    got = b.download(0, _abi.F_BV)
    mask = np.ones(got.shape, bool)
    mask[inner] = False
    assert np.array_equal(got[mask], ghosts_before[mask])  # ghost layers are not in the file: kept
    a.close()
    b.close()


def test_zones_are_matched_to_blocks_by_node_counts(tables, tmp_path):
    c1, c2 = box(tables, 12, 10, 8, seed=6), box(tables, 10, 8, 8, seed=7)
    a = Solver(c1.run, tables)
    a.add_block(c1)
    a.add_block(c2)
    a.initialize()
    path = tmp_path / "flowfield.plt"
    a.write_plt(path)
    b = Solver(c1.run, tables)
    b.add_block(box(tables, 10, 8, 8, seed=8))
    b.add_block(box(tables, 12, 10, 8, seed=9))
    print("plt block order", flush=True)
    b.read_plt(path)
    for blk_b, blk_a, c in ((0, 1, c2), (1, 0, c1)):
        g = c.ngg
        inner = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
        assert np.array_equal(b.download(blk_b, _abi.F_BV)[inner], a.download(blk_a, _abi.F_BV)[inner])
        assert np.array_equal(b.download(blk_b, _abi.F_SV)[inner], a.download(blk_a, _abi.F_SV)[inner])
    # a grid the file does not hold: an error (the reference would interpolate)
    c = Solver(c1.run, tables)
    c.add_block(box(tables, 12, 12, 8, seed=10))
    with pytest.raises(CoreflError, match="zones"):
        c.read_plt(path)
    for s in (a, b, c):
        s.close()


def test_missing_species_is_an_error(tables, tmp_path):
    case = box(tables, 12, 10, 8, seed=11)
    s = Solver(case.run, tables)
    s.add_block(case)
    path = tmp_path / "flowfield.plt"
    s.write_plt(path)
    raw = bytearray(path.read_bytes())
    # rename the variable "OH" (int32 per character) to "XH": the file no longer names that species
    code = lambda t: b"".join(struct.pack("<i", ord(ch)) for ch in t) + struct.pack("<i", 0)
    i = raw.find(code("OH"))
    assert i > 0
    raw[i:i + 4] = struct.pack("<i", ord("X"))
    bad = tmp_path / "renamed.plt"
    bad.write_bytes(bytes(raw))
    with pytest.raises(CoreflError, match="scalar 4 missing"):
        s.read_plt(bad)
    with pytest.raises(CoreflError, match="cannot open"):
        s.read_plt(tmp_path / "absent.plt")
    s.close()


def test_restart_matches_the_oracle_restarted_the_same_way(tables, tmp_path):
    """Three steps, output, restart (bv, sv from the file; cv, ghosts and properties from
    initialize), two more steps: against the oracle that takes the same three steps, starts a new
    run from its bv / sv interior and takes the same two steps."""
    dt = 2e-9
    case = box(tables, 16, 12, 10, seed=12)
    a = stepped(tables, case, 3, dt)
    a.write_output(tmp_path, step=3, solution_time=3 * dt)
    b = Solver(case.run, tables)
    b.add_block(box(tables, 16, 12, 10, seed=12))
    print("plt restart", flush=True)
    t, step = b.restart(tmp_path)
    assert (t, step) == (3 * dt, 3)
    for _ in range(2):
        b.step(dt)
    o = Oracle(case, h2_tables())
    o.initialize()
    for _ in range(3):
        o.step(dt)
    g = case.ngg
    inner = (slice(None), slice(g, -g), slice(g, -g), slice(g, -g))
    restart = box(tables, 16, 12, 10, seed=12)
    restart.bv[inner] = o.a["bv"][inner]
    restart.sv[inner] = o.a["sv"][inner]
    o2 = Oracle(restart, h2_tables())
    o2.initialize()
    for _ in range(2):
        o2.step(dt)
    for name, fid in (("cv", _abi.F_CV), ("bv", _abi.F_BV), ("sv", _abi.F_SV)):
        got = b.download(0, fid).reshape(o2.a[name].shape)[inner]
        ref = o2.a[name][inner]
        err = np.abs(got - ref).reshape(ref.shape[0], -1).max(axis=1) / np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
        assert np.all(err <= 1e-10), (name, err)
    a.close()
    b.close()
