"""The chemistry kernel compiled for the mechanism (chem_kernel 0, hipRTC)
against the table-driven kernel (chem_kernel 1) and the oracle
(src/FiniteRateChem.cu:1406-1681), on the Li-2004 mechanism (reaction types
1, 3, 4, 6) and on a variant of it that also holds the other rate forms:
irreversible (0), explicit backward rate only (2), a REV-given backward rate
on a third-body reaction (4), Lindemann (5) and four-parameter Troe (7).
The variant's values are arbitrary; they exercise the code, not the chemistry."""
import ctypes as C

import numpy as np
import pytest

from corefl_amd import Solver, _abi, load_tables, reacting_box_case
from oracle import Oracle, tables_from_dict

STAGE_TOL = 1e-11

# (reaction, field, value) changes that give the variant its rate forms
VARIANT = [
    (0, "reac_type", 0),
    (1, "reac_type", 2), (1, "rev_type", 1), (1, "A2", 3.1e12), (1, "b2", 0.3), (1, "Ea2", 4100.0),
    (4, "rev_type", 1), (4, "A2", 2.2e15), (4, "b2", -0.4), (4, "Ea2", 52000.0),
    (8, "reac_type", 7), (8, "troe_t2", 5182.0),
    (7, "reac_type", 5), (7, "A2", 1.0e18), (7, "b2", -1.0), (7, "Ea2", 0.0),
]


def variant_tables(base):
    """A copy of ChemTables with the VARIANT changes."""
    t = _abi.ChemTables()
    C.memmove(C.byref(t), C.byref(base), C.sizeof(t))
    for r, k, v in VARIANT:
        getattr(t, k)[r] = v
    return t


def variant_dict(d):
    d = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in d.items()}
    for r, k, v in VARIANT:
        d[k][r] = v
    return d


@pytest.fixture(scope="module")
def tables():
    return load_tables()


@pytest.mark.parametrize("variant", [False, True])
def test_generated_kernel_compiles(tables, variant):
    t = variant_tables(tables) if variant else tables
    log = C.create_string_buffer(1 << 16)
    st = _abi.load().corefl_chem_kernel_check(C.byref(t), log, len(log))
    assert st == 0, log.value.decode()


def test_check_rejects_bad_tables(tables):
    t = variant_tables(tables)
    t.n_spec = 0
    assert _abi.load().corefl_chem_kernel_check(C.byref(t), None, 0) == -1  # COREFL_ERR_ARG


def test_variant_holds_every_rate_form(tables):
    t = variant_tables(tables)
    assert sorted({t.reac_type[r] for r in range(t.n_reac)}) == [0, 1, 2, 3, 4, 5, 6, 7]
    assert any(t.rev_type[r] == 1 and t.reac_type[r] == 4 for r in range(t.n_reac))


def with_radicals(case, seed):
    """Adds up to 2% of every species to the H2/O2/N2 state (renormalised), so
    that every reaction runs in both directions."""
    rng = np.random.default_rng(seed)
    y = case.sv + 0.02 * rng.uniform(size=case.sv.shape)
    case.sv[:] = y / y.sum(axis=0, keepdims=True)
    return case


def normwise(a, b):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    return np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [False, True])
def test_chemistry_kernels_parity(tables, h2_tables_dict, variant):
    gpu_tables = variant_tables(tables) if variant else tables
    oc_tables = tables_from_dict(variant_dict(h2_tables_dict) if variant else h2_tables_dict)
    case = with_radicals(reacting_box_case(20, 16, 12, seed=15, tables=gpu_tables), seed=15)
    o = Oracle(case, oc_tables)
    o.initialize()
    o.op("zero_dq")
    o.op("chemistry")
    g = case.ngg
    res = {}
    for kernel in (0, 1):
        print(f"chemistry variant={variant} chem_kernel={kernel}", flush=True)
        case.run["chem_kernel"] = kernel
        s = Solver(case.run, gpu_tables)
        s.add_block(case)
        s.initialize()
        s.op(_abi.OP_ZERO_DQ)
        s.op(_abi.OP_CHEMISTRY)
        dq = s.download(0, _abi.F_DQ)
        rts = s.download(0, _abi.F_RTS).reshape(1, *case.padded_shape)[:, g:-g, g:-g, g:-g]
        s.close()
        err = normwise(dq, o.a["dq"])
        assert np.all(err <= STAGE_TOL), f"chem_kernel {kernel}: dq {err}"
        err = normwise(rts, o.a["rts"][None, g:-g, g:-g, g:-g])
        assert np.all(err <= STAGE_TOL), f"chem_kernel {kernel}: rts {err}"
        res[kernel] = dq
    # the two kernels evaluate the same expressions in the same order
    assert np.all(normwise(res[0], res[1]) <= 1e-13)
