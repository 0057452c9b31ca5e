"""XPBD and explicit integrators on the GPU vs the FP64 oracle."""
import numpy as np
import pytest

from conftest import beam_diag
from scenes import make_pair, max_err

pytestmark = pytest.mark.gpu

DIAG = beam_diag()
TOL1 = 1e-5 * DIAG
TOL100 = 1e-3 * DIAG

XPBD_CASES = [
    (3, (16, 4, 4), 0),   # P1 -> sqrt form (V0), mask-aware
    (7, (4, 2, 2), 2),    # P2 OptiShared -> energy form (V2)
    (9, (4, 2, 2), 2),    # Q2 OptiShared
    (8, (3, 1, 1), 2),    # P3 OptiShared
    (7, (4, 2, 2), 0),    # P2 with the XPBD version -> V0 for every element
    (6, (6, 2, 2), 1),    # Q1 Quadrature -> V2 arithmetic
]


@pytest.mark.parametrize("case", XPBD_CASES, ids=[f"el{c[0]}-v{c[2]}" for c in XPBD_CASES])
def test_xpbd_one_frame(engine, case):
    el, cells, version = case
    g, o, rest, _ = make_pair(engine, el, cells, solver=1, material=2, young=1e6, poisson=0.45,
                              xpbd_version=version)
    # one frame: dt = 0.01 with 50 substeps, as the XPBD P3/Q2 workload steps
    g.step(0.01, 50)
    o.step(0.01, 50)
    err = max_err(g.positions(), o.positions())
    disp = float(np.abs(o.positions() - rest).max())
    print(f"XPBD el={el} v={version} 1 frame err={err:.3e} disp={disp:.3e}", flush=True)
    assert disp > 1e-4
    assert err <= TOL1


def test_xpbd_many_frames(engine):
    g, o, rest, _ = make_pair(engine, 3, (16, 4, 4), solver=1, material=2, young=1e6, poisson=0.45,
                              xpbd_version=0)
    for _ in range(30):
        g.step(0.01, 20)
        o.step(0.01, 20)
    err = max_err(g.positions(), o.positions())
    print(f"XPBD P1 30 frames err={err:.3e}", flush=True)
    assert err <= TOL100


def test_xpbd_stable_neohooke_and_damping(engine):
    g, o, _, _ = make_pair(engine, 7, (4, 2, 2), solver=1, material=3, young=1e6, poisson=0.45,
                           global_damping=2.0, xpbd_version=2)
    for _ in range(3):
        g.step(0.01, 10)
        o.step(0.01, 10)
    assert max_err(g.positions(), o.positions()) <= TOL1 * 3


def test_xpbd_optigroup_high_order_is_inert(engine):
    """OptiGroup launches no projection for non-P1 elements (reference
    GPU_PBD_FEM.cu:671-692): only gravity and the constraints act."""
    g, o, _, _ = make_pair(engine, 7, (4, 2, 2), solver=1, material=2, young=1e6, poisson=0.45,
                           xpbd_version=3)
    g.step(0.01, 5)
    o.step(0.01, 5)
    assert max_err(g.positions(), o.positions()) <= TOL1


EXPLICIT_CASES = [(9, (4, 2, 2)), (8, (3, 1, 1)), (7, (4, 2, 2)), (3, (8, 2, 2))]


@pytest.mark.parametrize("case", EXPLICIT_CASES, ids=[f"el{c[0]}" for c in EXPLICIT_CASES])
def test_explicit_substeps(engine, case):
    el, cells = case
    # the explicit P3/Q2 workload: NeoHooke E=1e6 nu=0.45, +-1e-3 m noise, dt = 1e-5
    g, o, rest, _ = make_pair(engine, el, cells, solver=2, material=2, young=1e6, poisson=0.45, noise=1e-3, seed=7)
    g.step(2e-3, 200)      # 200 substeps of 1e-5
    o.step(2e-3, 200)
    err = max_err(g.positions(), o.positions())
    moved = max_err(o.positions(), rest)
    print(f"explicit el={el} 200 substeps err={err:.3e} moved={moved:.3e}", flush=True)
    assert moved > 1e-4
    assert err <= TOL1


def test_explicit_damping(engine):
    g, o, _, _ = make_pair(engine, 9, (4, 2, 2), solver=2, material=3, young=1e6, poisson=0.45, damping=1e-4,
                           noise=1e-3, seed=3)
    g.step(1e-3, 100)
    o.step(1e-3, 100)
    assert max_err(g.positions(), o.positions()) <= TOL1


FORCE_PASS_CASES = [(9, (4, 2, 2)), (8, (3, 1, 1)), (7, (4, 2, 2)), (6, (6, 2, 2)), (3, (8, 2, 2))]


@pytest.mark.parametrize("case", FORCE_PASS_CASES, ids=[f"el{c[0]}" for c in FORCE_PASS_CASES])
def test_explicit_element_pass_matches_the_gather(engine, case):
    """The element-centric force pass (the default, mode 2) and the per-vertex
    gather (mode 0) add the same terms in another order: both follow the oracle
    within TOL1 and each other within a hundredth of it."""
    import octopusengine_amd as oe

    el, cells = case
    kw = dict(solver=2, material=2, young=1e6, poisson=0.45, damping=1e-5, noise=1e-3, seed=5)
    g, o, rest, _ = make_pair(engine, el, cells, **kw)
    h, _, _, _ = make_pair(engine, el, cells, **kw)
    assert g.get_tile_ex()["mode"] == oe.PASS_TASKS
    h.set_tile(16, 1, oe.PASS_GATHER)
    assert h.get_tile_ex()["mode"] == oe.PASS_GATHER
    for s in (g, h, o):
        s.step(1e-3, 100)
    err = max_err(g.positions(), o.positions())
    diff = max_err(g.positions(), h.positions())
    moved = max_err(o.positions(), rest)
    print(f"explicit el={el} element pass err={err:.3e} vs gather {diff:.3e} moved={moved:.3e}", flush=True)
    assert moved > 1e-4
    assert err <= TOL1
    assert diff <= 1e-2 * TOL1


@pytest.mark.parametrize("solver", [1, 2], ids=["xpbd", "explicit"])
def test_graph_replays_match_eager_steps(engine, solver):
    """A scene stepped by HIP-graph replay (the element-centric explicit pass,
    or the XPBD colour launches) gives the same positions, bit for bit, as its
    eager twin after every replay, with host reads between replays."""
    kw = dict(solver=solver, material=2, young=1e6, poisson=0.45, noise=1e-3, seed=4)
    eager, o, _, _ = make_pair(engine, 9, (4, 2, 2), **kw)
    graph, _, _, _ = make_pair(engine, 9, (4, 2, 2), **kw)
    graph.set_graph(True)
    dt, sub = (0.01, 10) if solver == 1 else (2e-4, 20)
    for i in range(4):
        eager.step(dt, sub)
        graph.step(dt, sub)
        assert np.array_equal(eager.positions(), graph.positions()), f"replay {i}"
    for _ in range(4):
        o.step(dt, sub)
    assert max_err(graph.positions(), o.positions()) <= TOL1
