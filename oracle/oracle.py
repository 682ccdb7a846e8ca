"""ctypes binding of the oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker.  It mirrors the product's Solver interface
so that parity tests read the same on both sides.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
S, R = 20, 64
D, I = C.c_double, C.c_int


class OcTables(C.Structure):
    _fields_ = [
        ("n_spec", I), ("n_reac", I), ("nasa7or9", I),
        ("mw", D * S), ("imw", D * S), ("gas_const", D * S),
        ("t_low", D * S), ("t_mid", D * S), ("t_high", D * S),
        ("low_coeff", (D * 7) * S), ("high_coeff", (D * 7) * S),
        ("geometry", I * S),
        ("vis_coeff", D * S), ("LJ_potent_inv", D * S), ("ZRotF298", D * S),
        ("WjDivWi_to_One4th", (D * S) * S), ("sqrt_WiDivWjPl1Mul8", (D * S) * S),
        ("kb_over_eps_jk", (D * S) * S), ("binary_diffusivity_coeff", (D * S) * S),
        ("reac_type", I * R), ("rev_type", I * R), ("order", I * R),
        ("stoi_f", (I * S) * R), ("stoi_b", (I * S) * R),
        ("A", D * R), ("b", D * R), ("Ea", D * R), ("A2", D * R), ("b2", D * R), ("Ea2", D * R),
        ("third_body", (D * S) * R),
        ("troe_alpha", D * R), ("troe_t1", D * R), ("troe_t2", D * R), ("troe_t3", D * R),
        ("n_range", I * S), ("t_cut", (D * 5) * S), ("poly9", ((D * 9) * 4) * S),
        ("two_temperature", I), ("theta_v", D * S), ("n_elec", I * S),
        ("elec_theta", (D * 16) * S), ("elec_g", (D * 16) * S),
        ("lt_a", (D * S) * S), ("lt_b", (D * S) * S), ("park_sigma", D * S),
        ("trx_model", I), ("tcf_a", D * R), ("tcf_b", D * R), ("tcb_a", D * R), ("tcb_b", D * R),
    ]


class OcRun(C.Structure):
    _fields_ = [
        ("mixture", I), ("n_spec", I), ("n_var", I), ("n_scalar", I), ("n_scalar_transported", I),
        ("inviscid_scheme", I), ("viscous_order", I), ("reaction", I), ("gradp_in_diffusion", I),
        ("prandtl", D), ("cfl", D), ("weno_eps_scale", D),
        ("reconstruction", I), ("limiter", I), ("entropy_fix_factor", D),
        ("limit_flow", I), ("rho_inf", D), ("v_inf", D), ("p_inf", D),
        ("two_temperature", I), ("vt_relaxation", I),
        ("positive_preserving", I), ("pp_dt", D), ("shock_sensor", I), ("sensor_eps", D), ("sensor_threshold", D),
        ("chem_advance_method", I),
    ]


class OcBoundary(C.Structure):
    _fields_ = [("face", I), ("direction", I), ("type", I), ("range_start", I * 3), ("range_end", I * 3),
                ("thermal_type", I), ("inflow_sub", I), ("wall_temperature", D),
                ("density", D), ("u", D), ("v", D), ("w", D), ("pressure", D), ("temperature", D), ("mw", D),
                ("sv", D * S), ("temperature_ve", D)]


# conditions a label stands for when the setup defines none (Wall defaults:
# isothermal at 300 K, src/BoundCond.h:102-110)
DEFAULT_KIND = {2: dict(kind=2, thermal_type=0, wall_temperature=300.0), 3: dict(kind=3), 6: dict(kind=6),
                8: dict(kind=8)}


PD = C.POINTER(D)
PI = C.POINTER(I)


class OcInnerFace(C.Structure):
    _fields_ = [("range_start", I * 3), ("range_end", I * 3), ("face", I), ("direction", I),
                ("target_start", I * 3), ("target_face", I), ("target_direction", I),
                ("src_tar", I * 3), ("loop_dir", I * 3), ("target_loop_dir", I * 3), ("n_point", I * 3),
                ("target_block", I)]


class OcZone(C.Structure):
    _fields_ = [("mx", I), ("my", I), ("mz", I), ("ngg", I)] + \
        [(n, PD) for n in ("cv", "qn", "bv", "sv", "dq", "ff", "gf", "hf", "metric", "jac", "mul", "kappa", "gamma",
                           "acoustic", "mach", "rhoD", "rts", "dt_local", "tve", "kappa_ve", "shock_sensor")] + \
        [(n, PI) for n in ("btype_il", "btype_ir", "btype_jl", "btype_jr", "btype_kl", "btype_kr")] + \
        [("n_boundary", I), ("boundary", C.POINTER(OcBoundary)), ("n_inner", I), ("inner", C.POINTER(OcInnerFace)),
         ("cn", PD)]


_lib = None
_lib_fast = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load(fast: bool = False) -> C.CDLL:
    """The parity build, or with fast the -O3 -march=native build with FMA
    contraction (the CPU baseline only; built here on first use, for this
    machine's CPU)."""
    global _lib, _lib_fast
    if fast:
        if _lib_fast is None:
            # built per machine (and source version) outside the tree, so a copy built for another CPU never
            # travels with the repository
            import hashlib
            import tempfile
            cpu = Path("/proc/cpuinfo").read_text() if Path("/proc/cpuinfo").exists() else ""
            model = "".join(sorted({ln for ln in cpu.splitlines() if ln.startswith(("model name", "flags"))}))
            src = (HERE / "corefl_oracle.c").read_bytes()
            key = hashlib.sha1(model.encode() + src).hexdigest()[:16]
            out = Path(tempfile.gettempdir()) / f"corefl_oracle_fast_{key}.so"
            if not out.exists():
                tmp = out.with_suffix(f".{os.getpid()}.tmp")
                subprocess.run(["make", "-s", "-C", str(HERE), "fast", f"FAST_OUT={tmp}"], check=True)
                os.replace(tmp, out)
            _lib_fast = _bind(C.CDLL(str(out)))
        return _lib_fast
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = _bind(C.CDLL(str(LIB)))
    return _lib


def _bind(lib: C.CDLL) -> C.CDLL:
    if True:  # argument types of every entry point the tests and the baseline call
        zp, rp, tp = C.POINTER(OcZone), C.POINTER(OcRun), C.POINTER(OcTables)
        for name in ("oc_cv_from_bv", "oc_update_properties", "oc_viscous8", "oc_chemistry",
                     "oc_boundary_conditions", "oc_initialize"):
            getattr(lib, name).argtypes = [zp, rp, tp]
        lib.oc_zero_dq.argtypes = [zp, rp]
        lib.oc_eliminate_k_gradient.argtypes = [zp, rp]
        lib.oc_shock_sensor.argtypes = [zp, rp]
        lib.oc_inviscid_dir.argtypes = [zp, rp, tp, I]
        lib.oc_update_rk.argtypes = [zp, rp, tp, D, I]
        lib.oc_vt_relaxation.argtypes = [zp, rp, tp, D]
        lib.oc_local_time_step.argtypes = [zp, rp, tp]
        lib.oc_local_time_step.restype = D
        lib.oc_stage.argtypes = [zp, rp, tp, D, I]
        lib.oc_step.argtypes = [zp, rp, tp, D, I]
        lib.oc_step.restype = D
        lib.oc_enthalpy.argtypes = [tp, D, PD]
        lib.oc_cp.argtypes = [tp, D, PD]
        lib.oc_enthalpy_and_cp.argtypes = [tp, D, PD, PD]
        lib.oc_gibbs_div_rt.argtypes = [tp, D, PD]
        lib.oc_chem_source.argtypes = [tp, D, D, D, PD, PD, PD]
        lib.oc_transport.argtypes = [tp, D, D, PD, PD, PD, PD, PD]
        lib.oc_weno5.argtypes = [PD, PD, D, I]
        lib.oc_weno5.restype = D
        lib.oc_weno7.argtypes = [PD, PD, D, I]
        lib.oc_weno7.restype = D
        lib.oc_char_eigen.argtypes = [rp, tp, PD, D, PD, D, D, D, D, PD, PD]
        lib.oc_cv_reactor.argtypes = [tp, D, PD, PD, D, D, PD, PD, C.c_long, C.c_long]
        lib.oc_cv_reactor.restype = C.c_long
        lib.oc_cv_reactor_ex.argtypes = [tp, D, PD, PD, D, D, I, I]
        lib.oc_cv_reactor_ex.restype = C.c_long
        lib.oc_cv_reactor_rok4e.argtypes = [tp, D, PD, PD, D, D, I, D, D, C.POINTER(C.c_long)]
        lib.oc_cv_reactor_rok4e.restype = C.c_long
        zpp = C.POINTER(zp)
        lib.oc_register_inner.argtypes = [PI, PI, I, I, I, C.POINTER(OcInnerFace)]
        lib.oc_data_communication.argtypes = [zpp, I, rp, tp]
        lib.oc_viscous8_blocks.argtypes = [zpp, I, rp, tp]
        lib.oc_viscous2.argtypes = [zp, rp, tp]
        lib.oc_limit_flow.argtypes = [zp, rp, tp]
        lib.oc_initialize_blocks.argtypes = [zpp, I, rp, tp]
        lib.oc_stage_blocks.argtypes = [zpp, I, rp, tp, D, I]
        lib.oc_step_blocks.argtypes = [zpp, I, rp, tp, D, I]
        lib.oc_step_blocks.restype = D
        lib.oc_compute_cn_blocks.argtypes = [zpp, I, rp, tp]
        lib.oc_wu_reaction.argtypes = [zp, rp, tp, D]
        lib.oc_update_rk_wu.argtypes = [zp, rp, tp, D, I]
        lib.oc_wu_step_blocks.argtypes = [zpp, I, rp, tp, D, I]
        lib.oc_wu_step_blocks.restype = D
    return lib


def tables_from_dict(T: dict) -> OcTables:
    t = OcTables()
    t.n_spec, t.n_reac, t.nasa7or9 = T["n_spec"], T["n_reac"], int(T.get("nasa7or9", 7))
    ns, nr = T["n_spec"], T["n_reac"]
    for k in ("mw", "imw", "gas_const", "t_low", "t_mid", "t_high", "vis_coeff", "LJ_potent_inv", "ZRotF298"):
        getattr(t, k)[:ns] = [float(v) for v in T[k]]
    t.geometry[:ns] = [int(v) for v in T["geometry"]]
    for s in range(ns):
        t.low_coeff[s][:] = [float(v) for v in T["low_coeff"][s]]
        t.high_coeff[s][:] = [float(v) for v in T["high_coeff"][s]]
        for k in ("WjDivWi_to_One4th", "sqrt_WiDivWjPl1Mul8", "kb_over_eps_jk", "binary_diffusivity_coeff"):
            getattr(t, k)[s][:ns] = [float(v) for v in T[k][s]]
    for k in ("reac_type", "rev_type", "order"):
        getattr(t, k)[:nr] = [int(v) for v in T[k]]
    for k in ("A", "b", "Ea", "A2", "b2", "Ea2", "troe_alpha", "troe_t1", "troe_t2", "troe_t3"):
        getattr(t, k)[:nr] = [float(v) for v in T[k]]
    for r in range(nr):
        t.stoi_f[r][:ns] = [int(v) for v in T["stoi_f"][r]]
        t.stoi_b[r][:ns] = [int(v) for v in T["stoi_b"][r]]
        t.third_body[r][:ns] = [float(v) for v in T["third_body"][r]]
    if "n_range" in T:
        for s in range(ns):
            t.n_range[s] = int(T["n_range"][s])
            t.t_cut[s][:] = [float(v) for v in T["t_cut"][s]]
            for j in range(4):
                t.poly9[s][j][:] = [float(v) for v in T["poly9"][s][j]]
    if T.get("two_temperature"):
        t.two_temperature = 1
        for s in range(ns):
            t.theta_v[s] = float(T["theta_v"][s])
            t.park_sigma[s] = float(T["park_sigma"][s])
            t.n_elec[s] = int(T["n_elec"][s])
            t.elec_theta[s][:] = [float(v) for v in T["elec_theta"][s]]
            t.elec_g[s][:] = [float(v) for v in T["elec_g"][s]]
            t.lt_a[s][:ns] = [float(v) for v in T["lt_a"][s]]
            t.lt_b[s][:ns] = [float(v) for v in T["lt_b"][s]]
    t.trx_model = int(T.get("trx_model", 0))
    for k, default in (("tcf_a", 1.0), ("tcf_b", 0.0), ("tcb_a", 1.0), ("tcb_b", 0.0)):
        getattr(t, k)[:nr] = [float(v) for v in T.get(k, [default] * nr)]
    return t


def air5_tables(two_temperature: bool = True) -> OcTables:
    """AIR-5 (Park's 17 reactions), NASA-9, with or without the two-temperature data."""
    from chem_tables import build_tables  # noqa: E402  (oracle/ on sys.path)
    root = HERE.parent / "corefl-cpc_amd" / "data" / "air5_park"
    return tables_from_dict(build_tables(str(root / "mechanism.inp"), str(root / "therm.dat"), str(root / "tran.dat"),
                                         nasa=9, two_temperature=str(root / "two_temperature.dat")
                                         if two_temperature else None))


def h2_tables() -> OcTables:
    from chem_tables import build_tables  # noqa: E402  (oracle/ on sys.path)
    root = HERE.parent / "corefl-cpc_amd" / "data" / "h2_li2004"
    return tables_from_dict(build_tables(str(root / "mechanism.inp"), str(root / "therm.dat"),
                                         str(root / "tran.dat")))


def _p(a):
    return a.ctypes.data_as(PD)


def register_inner(src, tar, target_block: int, dim: int, ngg: int) -> OcInnerFace:
    """The oracle's restatement of InnerFace::register_boundary on one line pair."""
    f = OcInnerFace()
    load().oc_register_inner((I * 6)(*src), (I * 6)(*tar), target_block, dim, ngg, C.byref(f))
    return f


class _Block:
    """Arrays and oracle zone of one block (numpy, layout as the reference's DZone)."""

    def __init__(self, case, r):
        shp = case.padded_shape
        nv, nsc, ns = r["n_var"], max(1, r["n_scalar"]), max(1, r["n_spec"])
        isz = (case.mz, case.my, case.mx)
        z = lambda *s: np.zeros(s)
        self.a = dict(cv=z(nv, *shp), qn=z(nv, *shp), bv=np.ascontiguousarray(case.bv, dtype=float).copy(),
                      sv=np.zeros((nsc,) + shp), dq=z(nv, *isz), ff=z(nv, *shp), gf=z(nv, *shp), hf=z(nv, *shp),
                      metric=np.ascontiguousarray(case.metric, dtype=float).copy(),
                      jac=np.ascontiguousarray(case.jac, dtype=float).copy(), mul=z(*shp), kappa=z(*shp),
                      gamma=z(*shp), acoustic=z(*shp), mach=z(*shp), rhoD=z(ns, *shp), rts=z(*shp),
                      dt_local=z(*isz), tve=z(*shp), kappa_ve=z(*shp), shock_sensor=z(*shp), cn=z(nv, *isz))
        if r["n_scalar"]:
            self.a["sv"][:] = case.sv
        if case.tve is not None:
            self.a["tve"][:] = case.tve
        mx, my, mz = case.mx, case.my, case.mz
        bt = dict(il=np.zeros(my * mz, np.int32), ir=np.zeros(my * mz, np.int32), jl=np.zeros(mx * mz, np.int32),
                  jr=np.zeros(mx * mz, np.int32), kl=np.zeros(mx * my, np.int32), kr=np.zeros(mx * my, np.int32))
        kinds = {c["label"]: c["kind"] for c in case.conditions}
        for face, direction, typ, start, end in case.boundaries:
            if kinds.get(typ, DEFAULT_KIND.get(typ, {}).get("kind")) == 8:
                continue
            key = "ijk"[face] + ("l" if direction < 0 else "r")
            for k in range(start[2], end[2] + 1):
                for j in range(start[1], end[1] + 1):
                    for i in range(start[0], end[0] + 1):
                        idx = {0: j + k * my, 1: i + k * mx, 2: i + j * mx}[face]
                        bt[key][idx] = typ
        self.bt = bt
        self.faces = (OcBoundary * max(1, len(case.boundaries)))()
        conds = {c["label"]: c for c in case.conditions}
        for n, (face, direction, typ, start, end) in enumerate(case.boundaries):
            cond = conds.get(typ, DEFAULT_KIND.get(typ))
            if cond is None:
                raise ValueError(f"boundary label {typ} has no condition")
            f = self.faces[n]
            f.face, f.direction, f.type = face, direction, cond["kind"]
            for a in range(3):
                f.range_start[a] = start[a]
                f.range_end[a] = end[a]
            for k, v in cond.items():
                if k == "sv":
                    f.sv[:len(v)] = [float(x) for x in v]
                elif k not in ("label", "kind"):
                    setattr(f, k, v)
        dim = 2 if mz == 1 else 3
        self.inner = (OcInnerFace * max(1, len(case.inner)))()
        for n, (src, tar, tblk) in enumerate(case.inner):
            self.inner[n] = register_inner(src, tar, tblk, dim, case.ngg)
        zz = OcZone()
        zz.mx, zz.my, zz.mz, zz.ngg = mx, my, mz, case.ngg
        for k, v in self.a.items():
            setattr(zz, k, _p(v))
        for k, v in bt.items():
            setattr(zz, "btype_" + k, v.ctypes.data_as(PI))
        zz.n_boundary = len(case.boundaries)
        zz.boundary = self.faces
        zz.n_inner = len(case.inner)
        zz.inner = self.inner
        self.z = zz


class Oracle:
    """CPU restatement of one block or a list of blocks (1-to-1 interfaces)."""

    def __init__(self, case, tables: OcTables | None, fast: bool = False):
        self.lib = load(fast)
        cases = list(case) if isinstance(case, (list, tuple)) else [case]
        self.case = cases[0]
        r = cases[0].run
        self.run = OcRun(r["mixture"], r["n_spec"], r["n_var"], r["n_scalar"], r["n_scalar_transported"],
                         r["inviscid_scheme"], r["viscous_order"], r["reaction"], r["gradp_in_diffusion"],
                         r["prandtl"], r["cfl"], (r["rho_ref"] * r["v_ref"]) * (r["rho_ref"] * r["v_ref"]),
                         r.get("reconstruction", 2), r.get("limiter", 0), r.get("entropy_fix_factor", 0.125),
                         r.get("limit_flow", 0), r["rho_ref"], r["v_ref"], r.get("p_ref", 1.0),
                         r.get("two_temperature", 0), r.get("vt_relaxation", 0),
                         r.get("positive_preserving", 0), r["dt"] if r.get("fixed_time_step", 0) else 0.0,
                         r.get("shock_sensor", 0), r.get("sensor_eps", -1.0), r.get("sensor_threshold", 0.0),
                         r.get("chem_advance_method", 0))
        self.tables = tables if tables is not None else OcTables()
        self.blocks = [_Block(c, r) for c in cases]
        self.a = self.blocks[0].a
        self.z = self.blocks[0].z
        self.zones = (C.POINTER(OcZone) * len(self.blocks))(*[C.pointer(b.z) for b in self.blocks])
        self.nb = len(self.blocks)
        self.args = (C.byref(self.z), C.byref(self.run), C.byref(self.tables))
        self.margs = (self.zones, self.nb, C.byref(self.run), C.byref(self.tables))

    def initialize(self):
        self.lib.oc_initialize_blocks(*self.margs)

    def stage(self, rk, dt):
        self.lib.oc_stage_blocks(*self.margs, dt, rk)

    def step(self, dt=0.0, fixed=True):
        if self.case.run.get("temporal_scheme", 3) == 4:
            return self.lib.oc_wu_step_blocks(*self.margs, dt, 1 if fixed else 0)
        return self.lib.oc_step_blocks(*self.margs, dt, 1 if fixed else 0)

    def cfl_dt(self):
        return min(self.lib.oc_local_time_step(C.byref(b.z), C.byref(self.run), C.byref(self.tables))
                   for b in self.blocks)

    def op(self, name, rk=0, dt=0.0, blk=None):
        """One operator on one block (blk) or on all blocks in order (blk None)."""
        L = self.lib
        _, r, t = self.args
        if name == "viscous" and self.run.viscous_order == 8:
            if blk is not None and self.nb > 1:
                raise ValueError("the viscous operator spans all blocks")
            L.oc_viscous8_blocks(*self.margs)
            return
        if name == "halo":
            L.oc_data_communication(*self.margs)
            return
        for b in (self.blocks if blk is None else [self.blocks[blk]]):
            z = C.byref(b.z)
            if name == "zero_dq":
                L.oc_zero_dq(z, r)
            elif name == "chemistry":
                L.oc_chemistry(z, r, t)
            elif name == "viscous":
                if self.run.viscous_order == 2:
                    L.oc_viscous2(z, r, t)
            elif name.startswith("inviscid_"):
                L.oc_inviscid_dir(z, r, t, "xyz".index(name[-1]))
            elif name == "vt_relaxation":
                L.oc_vt_relaxation(z, r, t, dt)
            elif name == "update_rk":
                L.oc_update_rk(z, r, t, dt, rk)
            elif name == "bc":
                L.oc_boundary_conditions(z, r, t)
            elif name == "shock_sensor":
                L.oc_shock_sensor(z, r)
            elif name == "k_gradient":
                L.oc_eliminate_k_gradient(z, r)
            elif name == "properties":
                L.oc_update_properties(z, r, t)
            elif name == "cv_from_bv":
                L.oc_cv_from_bv(z, r, t)
            elif name == "limit_flow":
                L.oc_limit_flow(z, r, t)
            elif name == "copy_qn":
                b.a["qn"][:] = b.a["cv"]
            else:
                raise ValueError(name)
