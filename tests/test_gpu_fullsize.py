"""Size-independent properties at the benchmark's full sizes, where the FP64
oracle is too slow to run: the plain single scene (P2 64 x 16 x 16, 140,481
vertices), the Q2 32 x 8 x 8 beam, and the largest batched scene shape
(nx = 64, ny = nz = 16) of both types.

* Two independent formulations of a VBD colour pass - element tasks (partials
  in LDS, sum-factorised for Q2) and the per-vertex gather - agree after 1 step
  to 1e-5 of the bounding-box diagonal (the parity bound), and stay within
  1e-3 of it after 20 steps.
* Replaying a step from a HIP graph and launching it eagerly give identical
  positions; a one-scene batch ends within 1e-6 m of the scene stepped alone
  (the batched predictor and constraint kernels may round differently).
* A scene of the batched workload stepped in a batch beside other scenes ends
  within 1e-6 m of where it ends alone."""
import numpy as np
import pytest

import octopusengine_amd as oe

pytestmark = pytest.mark.gpu

SIZE = (1.0, 0.25, 0.25)
DIAG = float(np.linalg.norm(SIZE))
CASES = [(oe.TETRA10, (64, 16, 16)), (oe.HEXA27, (32, 8, 8))]
GATHER = {oe.TETRA10: (16, 2, oe.PASS_GATHER), oe.HEXA27: (8, 3, oe.PASS_GATHER)}


def scene(engine, el, cells, tile=None, young=5e5, poisson=0.49):
    s = oe.make_scene(engine, el, cells, SIZE, young=young, poisson=poisson, iterations=10)
    if tile:
        s.set_tile(*tile)
    return s


@pytest.mark.parametrize("el,cells", CASES)
def test_tasks_vs_gather(engine, el, cells):
    a = scene(engine, el, cells)                          # default: element tasks
    b = scene(engine, el, cells, tile=GATHER[el])
    assert a.get_tile()[2] == oe.PASS_TASKS
    a.step(0.005, 1)
    b.step(0.005, 1)
    e1 = float(np.abs(a.positions() - b.positions()).max())
    for _ in range(19):
        a.step(0.005, 1)
        b.step(0.005, 1)
    e20 = float(np.abs(a.positions() - b.positions()).max())
    moved = float(np.abs(a.positions() - a.init_positions()).max())
    print(f"el={el} {cells} N={a.n_vertices}: 1 step {e1:.2e}, 20 steps {e20:.2e} (moved {moved:.3f} m)", flush=True)
    assert moved > 1e-3 and np.isfinite(a.positions()).all()
    assert e1 <= 1e-5 * DIAG
    assert e20 <= 1e-3 * DIAG
    a.close()
    b.close()


@pytest.mark.parametrize("el,cells", CASES)
def test_graph_and_batch_identical(engine, el, cells):
    eager = scene(engine, el, cells)
    graph = scene(engine, el, cells)
    graph.set_graph(True)
    one = scene(engine, el, cells)
    b = oe.Batch(engine, [one])
    for _ in range(3):
        eager.step(0.005, 1)
        graph.step(0.005, 1)
        b.step(0.005, 1)
        engine.synchronize()
        assert np.array_equal(eager.positions(), graph.positions())
        assert float(np.abs(eager.positions() - one.positions()).max()) <= 1e-6
    for s in (eager, graph, one):
        s.close()
    b.close()


def test_largest_batched_shapes_in_a_mix(engine):
    """The largest scene shape of the batched workload (64 x 16 x 16) for both
    types, with smaller scenes of the other type around them."""
    specs = [(oe.TETRA10, (64, 16, 16), 5e5, 0.49), (oe.HEXA27, (64, 16, 16), 1e6, 0.45),
             (oe.HEXA27, (8, 4, 4), 1e7, 0.3), (oe.TETRA10, (12, 5, 7), 1e5, 0.4)]
    mixed = [scene(engine, el, c, young=E, poisson=nu) for el, c, E, nu in specs]
    alone = [scene(engine, el, c, young=E, poisson=nu) for el, c, E, nu in specs]
    b = oe.Batch(engine, mixed)
    for _ in range(2):
        b.step(0.005, 1)
        for s in alone:
            s.step(0.005, 1)
    for m, a in zip(mixed, alone):
        err = float(np.abs(m.positions() - a.positions()).max())
        print(f"N={m.n_vertices}: |batch - alone| = {err:.2e}", flush=True)
        assert err <= 1e-6, m.n_vertices
    for s in mixed + alone:
        s.close()
    b.close()
