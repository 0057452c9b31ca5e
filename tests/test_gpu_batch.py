"""Batched stepping: every scene of a mixed batch ends where it ends when
stepped alone, and where the FP64 oracle puts it."""
import numpy as np
import pytest

import octopusengine_amd as oe
from conftest import beam_diag
from scenes import make_pair, max_err

pytestmark = pytest.mark.gpu

TOL1 = 1e-5 * beam_diag()

MIX = [
    dict(el=oe.TETRA10, cells=(6, 2, 2), young=5e5, poisson=0.49),
    dict(el=oe.HEXA27, cells=(4, 2, 2), young=1e6, poisson=0.45),
    dict(el=oe.TETRA10, cells=(3, 1, 2), young=2e6, poisson=0.3),
    dict(el=oe.HEXA27, cells=(2, 1, 1), young=3e5, poisson=0.49),
    dict(el=oe.TETRA, cells=(8, 2, 2), young=1e6, poisson=0.4),
    dict(el=oe.TETRA20, cells=(2, 1, 1), young=1e6, poisson=0.45),
    dict(el=oe.HEXA, cells=(4, 2, 1), young=5e5, poisson=0.45),
]


def _make(engine, spec, **kw):
    g, o, rest, _ = make_pair(engine, spec["el"], spec["cells"], young=spec["young"], poisson=spec["poisson"], **kw)
    return g, o


def test_batch_equals_alone_and_oracle(engine):
    batched = [_make(engine, s) for s in MIX]
    alone = [_make(engine, s)[0] for s in MIX]
    b = oe.Batch(engine, [g for g, _ in batched])
    steps = 3
    for _ in range(steps):
        b.step(0.005, 1)
        for a in alone:
            a.step(0.005, 1)
        for _, o in batched:
            o.step(0.005, 1)
    for (g, o), a, spec in zip(batched, alone, MIX):
        pb, pa = g.positions(), a.positions()
        e_alone = max_err(pb, pa)
        e_orc = max_err(pb, o.positions())
        print(f"batch el={spec['el']} cells={spec['cells']} |batch-alone|={e_alone:.3e} |batch-oracle|={e_orc:.3e}",
              flush=True)
        assert e_alone <= 1e-6
        assert e_orc <= 3 * TOL1
    b.close()


def test_batch_tiles_with_gradient_cache(engine):
    """Per-element-type tiles, with and without the gradient cache, give each
    scene the step it gets alone with the same tile."""
    tiles = {oe.TETRA10: (16, 1, oe.PASS_TASKS), oe.HEXA27: (32, 1, oe.PASS_TASKS), oe.TETRA: (8, 1, oe.PASS_CACHE),
             oe.TETRA20: (64, 1, oe.PASS_TASKS), oe.HEXA: (16, 1, oe.PASS_CACHE)}
    batched = [_make(engine, s)[0] for s in MIX]
    alone = [_make(engine, s, tile=tiles[s["el"]])[0] for s in MIX]
    b = oe.Batch(engine, batched)
    for el, t in tiles.items():
        b.set_tile(el, *t)
    for _ in range(2):
        b.step(0.005, 1)
        for a in alone:
            a.step(0.005, 1)
    for g, a, spec in zip(batched, alone, MIX):
        err = max_err(g.positions(), a.positions())
        print(f"tiles el={spec['el']} |batch-alone|={err:.3e}", flush=True)
        assert err <= 1e-6
    b.close()


def test_batch_graph_replay_matches_eager(engine):
    s1 = [_make(engine, s)[0] for s in MIX[:4]]
    s2 = [_make(engine, s)[0] for s in MIX[:4]]
    b1 = oe.Batch(engine, s1)
    b2 = oe.Batch(engine, s2)
    b2.set_graph(True)
    for i in range(4):
        b1.step(0.005, 1)
        b2.step(0.005, 1)
        engine.synchronize()
        for a, c in zip(s1, s2):
            assert max_err(a.positions(), c.positions()) == 0.0, f"replay {i}"


def test_batch_streams_give_the_same_steps(engine):
    """One stream, or the Q2 colour passes on a second stream beside the
    others (eager and graph replay): the same positions, bit for bit."""
    s1 = [_make(engine, s)[0] for s in MIX]
    s2 = [_make(engine, s)[0] for s in MIX]
    s3 = [_make(engine, s)[0] for s in MIX]
    b1, b2, b3 = oe.Batch(engine, s1), oe.Batch(engine, s2), oe.Batch(engine, s3)
    b1.set_streams(1)
    b3.set_graph(True)   # two streams (the default), captured
    with pytest.raises(oe.OEError):
        b2.set_streams(3)
    for i in range(3):
        for b in (b1, b2, b3):
            b.step(0.005, 1)
        engine.synchronize()
        for a, c, d in zip(s1, s2, s3):
            assert max_err(a.positions(), c.positions()) == 0.0, f"step {i}: two streams, eager"
            assert max_err(a.positions(), d.positions()) == 0.0, f"step {i}: two streams, graph"
    for b in (b1, b2, b3):
        b.close()


def test_batch_tiles_and_timing(engine):
    scenes = [_make(engine, s)[0] for s in MIX[:2]]
    ref = [_make(engine, s)[0] for s in MIX[:2]]
    b = oe.Batch(engine, scenes)
    b.set_tile(oe.TETRA10, 8, 4)
    b.set_tile(oe.HEXA27, 32, 3)
    b.set_timing(True)
    b.step(0.005, 2)
    ms, n = b.get_timing()
    assert n > 0 and ms > 0
    for r in ref:
        r.step(0.005, 2)
    for a, r in zip(scenes, ref):
        assert max_err(a.positions(), r.positions()) <= 1e-6
    info = b.info()
    assert info["vertex_updates"] == sum(s.n_vertices * 10 for s in scenes)


def test_batch_rejects_unsupported(engine):
    xyz, topo = oe.beam_mesh(oe.TETRA10, (2, 1, 1))
    x = oe.Scene(engine, oe.TETRA10, xyz, topo, solver=oe.SOLVER_XPBD)
    with pytest.raises(oe.OEError):
        oe.Batch(engine, [x])
    c = oe.Scene(engine, oe.TETRA10, xyz, topo, rho=0.9)
    with pytest.raises(oe.OEError):
        oe.Batch(engine, [c])
