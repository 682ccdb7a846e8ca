"""C ABI checks that need no GPU: the library loads, exports every function
include/corefl_amd.h declares, and its host-side readers (Chemkin tables,
setup.txt) agree with the oracle's independent reader."""
import ctypes as C
import re
import shutil
from pathlib import Path

import numpy as np
import pytest

from corefl_amd import _abi, load_tables
from chem_tables import build_tables

REPO = Path(__file__).resolve().parent.parent


def header_functions():
    text = (REPO / "include" / "corefl_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(corefl_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_abi.EXPORTS) == names


AIR5 = REPO / "corefl-cpc_amd" / "data" / "air5_park"


def compare_tables(P, T):
    assert (P.n_spec, P.n_reac) == (T["n_spec"], T["n_reac"])
    for k, v in T.items():
        if k in ("species", "n_spec", "n_reac"):
            continue
        v = np.asarray(v)
        pv = np.array(getattr(P, k))
        if v.ndim:
            pv = pv[tuple(slice(0, n) for n in v.shape)]
        assert np.array_equal(pv, v), k


def test_chem_tables_match_oracle_reader(h2_tables_dict):
    compare_tables(load_tables(), h2_tables_dict)


def test_nasa9_two_temperature_tables_match_oracle_reader():
    """AIR-5 (N2, O2, NO, N, O; Park's 17 reactions) with NASA-9 thermodynamics
    and the two-temperature metadata, both readers."""
    files = [str(AIR5 / f) for f in ("mechanism.inp", "therm.dat", "tran.dat")]
    P = _abi.ChemTables()
    st = _abi.load().corefl_chem_load_ex(*[f.encode() for f in files], 9,
                                         str(AIR5 / "two_temperature.dat").encode(), C.byref(P))
    assert st == 0
    T = build_tables(*files, nasa=9, two_temperature=str(AIR5 / "two_temperature.dat"))
    compare_tables(P, T)
    assert P.nasa7or9 == 9 and P.two_temperature == 1 and P.trx_model == 1
    assert [P.n_range[i] for i in range(5)] == [3] * 5
    # N2: 3395 K vibrational temperature, 15 electronic levels, Ea of N2 dissociation 113200 K
    assert P.theta_v[0] == 3395.0 and P.n_elec[0] == 15 and abs(P.Ea[0] - 113200.0) < 1e-6


def test_chem_load_reports_missing_files(tmp_path):
    lib = _abi.load()
    t = _abi.ChemTables()
    st = lib.corefl_chem_load(str(tmp_path / "nope.inp").encode(), b"x", b"y", C.byref(t))
    assert st == -3  # COREFL_ERR_IO


def test_setup_reader_reference_format(tmp_path):
    """A reference-style setup.txt (types, arrays, structs, comments) is read
    into the run switches; defaults follow src/Parameter.cpp:358-546."""
    case = tmp_path / "case"
    chem = case / "input" / "chemistry"
    chem.mkdir(parents=True)
    src = REPO / "corefl-cpc_amd" / "data" / "h2_li2004"
    for f in ("mechanism.inp", "therm.dat", "tran.dat"):
        shutil.copy(src / f, chem / f)
    (case / "input" / "setup.txt").write_text("""// a reference-format setup file
int species = 1
int reaction = 1
string mechanism_file = chemistry/mechanism.inp
string therm_file = chemistry/therm.dat
string transport_file = chemistry/tran.dat
int therm_nasa_7_9 = 7
int inviscid_scheme = 72
int viscous_order = 8
bool fixed_time_step = 1
real dt = 2.5e-9
array string boundary_conditions {
    periodic
}
struct periodic {
    string type = periodic
    int label = 8
}
string reference_state = farstream
struct farstream {
    string type = inflow
    int label = 5
    real temperature = 1000
    real pressure = 101325
    real velocity = 50
    real O2 = 0.233
    real N2 = 0.767
}
""")
    lib = _abi.load()
    run, tab = _abi.RunDesc(), _abi.ChemTables()
    assert lib.corefl_setup_load(str(case).encode(), C.byref(run), C.byref(tab)) == 0
    assert (run.mixture, run.n_spec, run.n_var, run.n_scalar) == (1, 9, 14, 9)
    assert (run.inviscid_scheme, run.viscous_order, run.reaction, run.fixed_time_step) == (72, 8, 1, 1)
    assert run.dt == 2.5e-9 and run.cfl == 0.8 and run.prandtl == 0.72
    mw_air = 1 / (0.233 / tab.mw[2] + 0.767 / tab.mw[8])
    assert run.rho_ref == pytest.approx(101325 * mw_air / (8314.462618 * 1000), rel=1e-12)
    assert run.v_ref == 50.0
    assert tab.n_reac == 19


def test_setup_reader_two_temperature(tmp_path):
    """A two-temperature AIR-5 setup (NASA-9 thermo, the two_temperature_file key, Mach-number freestream):
    n_var = 6 + n_spec with E_ve last, V-T relaxation on by default (src/Parameter.cpp:424-426), the
    control-temperature model on for the 17-reaction mechanism, and the reference velocity from the Mach
    number with the NASA-9 cp."""
    case = tmp_path / "case"
    chem = case / "input" / "chemistry"
    chem.mkdir(parents=True)
    for f in ("mechanism.inp", "therm.dat", "tran.dat", "two_temperature.dat"):
        shutil.copy(AIR5 / f, chem / f)
    (case / "input" / "setup.txt").write_text("""int species = 1
int reaction = 1
string mechanism_file = chemistry/mechanism.inp
string therm_file = chemistry/therm.dat
string transport_file = chemistry/tran.dat
int therm_nasa_7_9 = 9
string two_temperature_file = chemistry/two_temperature.dat
int inviscid_scheme = 52
int viscous_order = 2
string reference_state = freestream
struct freestream {
    string type = inflow
    int label = 5
    real mach = 10.0
    real temperature = 220.0
    real pressure = 1000.0
    real N2 = 0.767
    real O2 = 0.233
}
""")
    run, tab = _abi.RunDesc(), _abi.ChemTables()
    assert _abi.load().corefl_setup_load(str(case).encode(), C.byref(run), C.byref(tab)) == 0
    assert (run.mixture, run.n_spec, run.n_var, run.n_scalar, run.n_scalar_transported) == (1, 5, 11, 6, 6)
    assert (run.two_temperature, run.vt_relaxation) == (1, 1)
    assert (tab.nasa7or9, tab.two_temperature, tab.trx_model, tab.n_reac) == (9, 1, 1, 17)
    mw = 1 / (0.767 / tab.mw[0] + 0.233 / tab.mw[1])
    R = 8314.462618 / mw
    # cp of cold air at 220 K is close to 3.5 R: gamma within 0.5% of 1.4
    c = run.v_ref / 10.0
    assert abs(c * c / (R * 220.0) - 1.4) < 0.007


STRUCTS = {"ChemTables": "corefl_chem_tables", "RunDesc": "corefl_run_desc", "Boundary": "corefl_boundary",
           "BcCondition": "corefl_bc_condition", "InnerFace": "corefl_inner_face", "BlockDesc": "corefl_block_desc",
           "ReactorRequest": "corefl_reactor_request", "ReactorResult": "corefl_reactor_result",
           "TubeRequest": "corefl_tube_request", "TubeResult": "corefl_tube_result"}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_the_header(tmp_path):
    """Size and every field offset of the ctypes mirrors equal the C compiler's
    for the header's structs (a field added on one side only fails here)."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "corefl_amd.h"', "int main(void) {"]
    for py, c in STRUCTS.items():
        lines.append(f'  printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(_abi, py)._fields_:
            lines.append(f'  printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", f"-I{REPO / 'include'}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            py, name, v = ln.split()
            got[(py, name)] = int(v)
    for py in STRUCTS:
        cls = getattr(_abi, py)
        assert got[(py, "size")] == C.sizeof(cls), py
        for name, _ in cls._fields_:
            assert got[(py, name)] == getattr(cls, name).offset, (py, name)
