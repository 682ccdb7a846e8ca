"""Binding of the reactor request server (corefl_serve_* in include/corefl_amd.h).

Requests are 0-D constant-volume reactors; the server keeps them in device
slots, admits queued requests as slots free up and advances every active
slot with one launch per progress() call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi

SCHEMES = {"point-implicit": 0, "rok4e": 1, "rk3": 3}


class ServeError(RuntimeError):
    pass


def h2_air_mass_fractions(phi: float, mw: np.ndarray) -> np.ndarray:
    """H2/air at equivalence ratio phi (2 phi H2 + O2 + 3.76 N2) as mass
    fractions in the Li-2004 species order (H2 H O2 O OH HO2 H2O2 H2O N2)."""
    x = np.zeros(9)
    x[0], x[2], x[8] = 2.0 * phi, 1.0, 3.76
    x /= x.sum()
    y = x * mw[:9]
    return y / y.sum()


class ReactorServer:
    def __init__(self, tables: _abi.ChemTables, budget_bytes: int = 64 << 20, steps_per_launch: int = 64):
        self.lib = _abi.load()
        self.tables = tables
        self.h = C.c_void_p()
        st = self.lib.corefl_serve_open(C.byref(tables), budget_bytes, steps_per_launch, C.byref(self.h))
        if st != 0:
            msg = self.lib.corefl_serve_error(self.h).decode() if self.h else str(st)
            self.close()
            raise ServeError(f"corefl_serve_open failed ({st}): {msg}")

    @property
    def capacity(self) -> int:
        return self.lib.corefl_serve_capacity(self.h)

    def submit(self, temperature, pressure, y, t_end, dt, scheme="rok4e", adaptive=True, rtol=1e-4, atol=1e-8,
               conserve_mass=False, max_steps=0) -> int:
        q = _request(temperature, pressure, y, t_end, dt, scheme, adaptive, rtol, atol, conserve_mass, max_steps)
        t = self.lib.corefl_serve_submit(self.h, C.byref(q))
        if t < 0:
            raise ServeError(f"corefl_serve_submit failed ({t}): {self.lib.corefl_serve_error(self.h).decode()}")
        return t

    def progress(self) -> int:
        pending = C.c_int(0)
        st = self.lib.corefl_serve_progress(self.h, C.byref(pending))
        if st != 0:
            raise ServeError(f"corefl_serve_progress failed ({st}): {self.lib.corefl_serve_error(self.h).decode()}")
        return pending.value

    def poll(self, ticket: int):
        r = _abi.ReactorResult()
        st = self.lib.corefl_serve_poll(self.h, ticket, C.byref(r))
        if st < 0:
            raise ServeError(f"corefl_serve_poll failed ({st}): {self.lib.corefl_serve_error(self.h).decode()}")
        if st == 0:
            return None
        return self._result(r)

    def _result(self, r):
        """status 0: reached t_end; -5 (COREFL_ERR_STATE): failed for `reason`
        (1 non-finite state, 2 step collapse, 3 step cap)."""
        ns = self.tables.n_spec
        return dict(temperature=r.temperature, pressure=r.pressure, time=r.time, y=np.array(r.y[:ns]),
                    steps=r.steps, rejected=r.rejected, status=r.status, reason=r.reason)

    def drain(self, max_results: int = 4096):
        """Finished results as {ticket: result}, removed from the server."""
        t = (C.c_long * max_results)()
        r = (_abi.ReactorResult * max_results)()
        n = self.lib.corefl_serve_drain(self.h, t, r, max_results)
        if n < 0:
            raise ServeError(f"corefl_serve_drain failed ({n})")
        return {t[i]: self._result(r[i]) for i in range(n)}

    def run_until_done(self, tickets):
        out = {}
        while True:
            pending = self.progress()
            for t in tickets:
                if t not in out:
                    r = self.poll(t)
                    if r is not None:
                        out[t] = r
            if pending == 0:
                break
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.corefl_serve_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BC_OUTFLOW, BC_SYMMETRY = 6, 3
TUBE_RUN = dict(mixture=1, n_spec=9, n_var=14, n_scalar=9, n_scalar_transported=9, inviscid_scheme=72,
                viscous_order=0, reaction=1, fixed_time_step=0, dt=0.0, cfl=0.5, prandtl=0.72, rho_ref=1.0,
                v_ref=400.0)


class TubeServer:
    """Batched 1-D reactive shock tubes (corefl_tubes_*): all tubes in flight
    advance together, each on its own CFL step, in one padded device row."""

    def __init__(self, tables: _abi.ChemTables, run: dict | None = None, budget_bytes: int = 256 << 20,
                 steps_per_progress: int = 16):
        from .solver import run_desc
        self.lib = _abi.load()
        self.tables = tables
        self.run = {**TUBE_RUN, **(run or {})}
        self.h = C.c_void_p()
        st = self.lib.corefl_tubes_open(C.byref(run_desc(self.run)), C.byref(tables), budget_bytes,
                                        steps_per_progress, C.byref(self.h))
        if st != 0:
            msg = self.lib.corefl_tubes_error(self.h).decode() if self.h else str(st)
            self.close()
            raise ServeError(f"corefl_tubes_open failed ({st}): {msg}")
        self.sizes = {}

    @property
    def capacity(self) -> int:
        return self.lib.corefl_tubes_capacity(self.h)

    def submit(self, n_cells, length, t_end, left, right, n_left=None, bc=("outflow", "outflow"),
               max_steps=0) -> int:
        """left / right: (T, p, u, Y) of the two initial states; max_steps > 0
        fails a tube still short of t_end after that many steps."""
        q = _abi.TubeRequest()
        q.max_steps = int(max_steps)
        q.n_cells, q.length, q.t_end = int(n_cells), float(length), float(t_end)
        q.n_left = int(n_cells // 2 if n_left is None else n_left)
        kinds = {"outflow": BC_OUTFLOW, "symmetry": BC_SYMMETRY, "wall": BC_SYMMETRY}
        q.bc_left, q.bc_right = (kinds[b] if isinstance(b, str) else int(b) for b in bc)
        for side, (T, p, u, y) in enumerate((left, right)):
            q.temperature[side], q.pressure[side], q.velocity[side] = T, p, u
            y = np.asarray(y, dtype=float)
            for l, v in enumerate(y):
                q.y[side][l] = float(v)
        t = self.lib.corefl_tubes_submit(self.h, C.byref(q))
        if t < 0:
            raise ServeError(f"corefl_tubes_submit failed ({t}): {self.lib.corefl_tubes_error(self.h).decode()}")
        self.sizes[t] = int(n_cells)
        return t

    def progress(self) -> int:
        pending = C.c_int(0)
        st = self.lib.corefl_tubes_progress(self.h, C.byref(pending))
        if st != 0:
            raise ServeError(f"corefl_tubes_progress failed ({st}): {self.lib.corefl_tubes_error(self.h).decode()}")
        return pending.value

    def poll(self, ticket: int):
        """None while queued or in flight; else dict(time, steps, profile) with
        profile[(6 + n_spec), n_cells] = rho, u, v, w, p, T, Y_l."""
        n = self.sizes.get(ticket, 0)
        rows = 6 + self.tables.n_spec
        prof = np.zeros(rows * max(n, 1))
        r = _abi.TubeResult()
        st = self.lib.corefl_tubes_poll(self.h, ticket, C.byref(r), prof.ctypes.data_as(C.POINTER(C.c_double)),
                                        prof.size)
        if st < 0:
            raise ServeError(f"corefl_tubes_poll failed ({st}): {self.lib.corefl_tubes_error(self.h).decode()}")
        if st == 0:
            return None
        self.sizes.pop(ticket, None)
        return dict(time=r.time, steps=r.steps, status=r.status, reason=r.reason,
                    profile=prof[:rows * r.n_cells].reshape(rows, r.n_cells))

    def ready(self, max_tickets: int = 4096):
        """Finished tickets not yet polled."""
        t = (C.c_long * max_tickets)()
        n = self.lib.corefl_tubes_ready(self.h, t, max_tickets)
        if n < 0:
            raise ServeError(f"corefl_tubes_ready failed ({n})")
        return [t[i] for i in range(n)]

    def run_until_done(self, tickets):
        out, want = {}, set(tickets)
        while True:
            pending = self.progress()
            for t in self.ready():
                if t in want:
                    out[t] = self.poll(t)
            if pending == 0:
                break
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.corefl_tubes_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NOALLOC_MAX_REACTORS = 4096  # COREFL_NOALLOC_MAX_REACTORS


def _request(temperature, pressure, y, t_end, dt, scheme="rok4e", adaptive=True, rtol=1e-4, atol=1e-8,
             conserve_mass=False, max_steps=0) -> _abi.ReactorRequest:
    q = _abi.ReactorRequest()
    q.temperature, q.pressure, q.t_end, q.dt = temperature, pressure, t_end, dt
    y = np.asarray(y, dtype=float)
    q.y[:len(y)] = [float(v) for v in y]
    q.scheme = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
    q.adaptive, q.rtol, q.atol, q.conserve_mass = int(adaptive), rtol, atol, int(conserve_mass)
    q.max_steps = int(max_steps)
    return q


def run_noalloc(tables: _abi.ChemTables, requests, steps_per_launch: int = 64) -> list:
    """One batch of reactors through corefl_serve_run_noalloc (no device
    allocation): requests are dicts of ReactorServer.submit's arguments; the
    results come back in request order, as ReactorServer's."""
    lib = _abi.load()
    n = len(requests)
    req = (_abi.ReactorRequest * max(1, n))()
    for i, q in enumerate(requests):
        req[i] = _request(**q)
    res = (_abi.ReactorResult * max(1, n))()
    st = lib.corefl_serve_run_noalloc(C.byref(tables), req, n, steps_per_launch, res)
    if st != 0:
        raise ServeError(f"corefl_serve_run_noalloc failed ({st})")
    ns = tables.n_spec
    return [dict(temperature=r.temperature, pressure=r.pressure, time=r.time, y=np.array(r.y[:ns]),
                 steps=r.steps, rejected=r.rejected, status=r.status, reason=r.reason) for r in res[:n]]
