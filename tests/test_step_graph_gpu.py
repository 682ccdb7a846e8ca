"""The recorded step (run.step_graph): one HIP-graph replay per RK3 step with
the blocks' work on concurrent streams and dt in device memory.  Every replay
must give the same state as the step issued operator by operator: the same
kernels run on the same data, only their launch order across independent
blocks differs.  Cases: the 9-species single block (8th-order viscous terms,
chemistry, fixed dt), twelve Air blocks with inner faces and the CFL step,
the 9-species multi-block case (inner-face viscous-flux exchange) and the
two-temperature box (V-T relaxation reads the device dt); the fixed-dt graph
is rebuilt when dt changes, and replays alternate with eager work."""
import numpy as np
import pytest

from corefl_amd import (Solver, _abi, air5_box_case, load_air5_tables, load_tables, multiblock_case,
                        reacting_box_case)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h2():
    return load_tables()


def solvers(cases, tables):
    cases = cases if isinstance(cases, list) else [cases]
    out = []
    for graph in (0, 1):
        r = dict(cases[0].run)
        r["step_graph"] = graph
        s = Solver(r, tables)
        for c in cases:
            s.add_block(c)
        s.initialize()
        out.append(s)
    print(f"{len(cases)} block(s), mixture {cases[0].run['mixture']}, fixed {cases[0].run['fixed_time_step']}",
          flush=True)
    return out


def same_state(a, b, nblk, fields=(_abi.F_CV, _abi.F_BV, _abi.F_SV), tol=0.0):
    for blk in range(nblk):
        for f in fields:
            x, y = a.download(blk, f), b.download(blk, f)
            scale = max(np.abs(x).max(), 1e-300)
            assert np.abs(x - y).max() <= tol * scale, f"block {blk} field {f}: {np.abs(x - y).max() / scale}"


def test_graph_single_block_reacting(h2):
    case = reacting_box_case(24, 16, 12, seed=41, tables=h2)
    eager, graph = solvers(case, h2)
    dt = 0.5 * eager.cfl_dt()
    for n in range(4):
        d = dt if n < 2 else 0.7 * dt   # the graph is rebuilt for the new dt
        eager.step(d)
        graph.step(d)
        same_state(eager, graph, 1)
    eager.close()
    graph.close()


def test_graph_multiblock_air_cfl():
    blocks = multiblock_case(split=(3, 2, 2), seed=7, physics="air", rotate=True)
    eager, graph = solvers(blocks, None)
    for _ in range(5):
        d1 = eager.step()
        d2 = graph.step()
        assert d1 == d2
        same_state(eager, graph, len(blocks), fields=(_abi.F_CV, _abi.F_BV))
        # eager work between replays: a properties pass on both
        eager.op(_abi.OP_PROPERTIES)
        graph.op(_abi.OP_PROPERTIES)
    eager.close()
    graph.close()


def test_graph_multiblock_reacting(h2):
    blocks = multiblock_case(split=(2, 2, 1), seed=9, physics="h2", rotate=True)
    for c in blocks:
        c.run.update(reaction=1, viscous_order=8)
    eager, graph = solvers(blocks, h2)
    dt = 0.3 * eager.cfl_dt()
    for _ in range(3):
        eager.step(dt)
        graph.step(dt)
        same_state(eager, graph, len(blocks))
    eager.close()
    graph.close()


def test_graph_two_temperature():
    tb = load_air5_tables(two_temperature=True)
    case = air5_box_case(16, 12, 10, seed=42, tables=tb, two_temperature=True)
    eager, graph = solvers(case, tb)
    dt = 0.5 * eager.cfl_dt()
    for _ in range(3):
        eager.step(dt)
        graph.step(dt)
        same_state(eager, graph, 1, fields=(_abi.F_CV, _abi.F_BV, _abi.F_SV, _abi.F_TVE))
    eager.close()
    graph.close()
