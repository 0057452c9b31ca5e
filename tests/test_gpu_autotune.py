"""Tile autotuner: the tuned layouts leave the scenes' state untouched, are
cached per (device, element type, size bucket, form), and step within the
oracle tolerance (scenes) or bit-for-bit like the scenes stepped alone
(batches, to 1e-6 m)."""
import numpy as np
import pytest

import octopusengine_amd as oe
from conftest import beam_diag
from scenes import make_pair, max_err

pytestmark = pytest.mark.gpu

TOL1 = 1e-5 * beam_diag()


@pytest.mark.parametrize("el,cells", [(7, (8, 4, 4)), (9, (6, 3, 3)), (3, (16, 4, 4))])
def test_scene_autotune_then_step_matches_oracle(engine, el, cells):
    oe.autotune_clear()
    g, o, _, _ = make_pair(engine, el, cells, young=5e5, poisson=0.49)
    g.step(0.005, 1)
    o.step(0.005, 1)
    before = (g.positions().copy(), g.last_positions().copy(), g.velocities().copy())
    t = g.autotune()
    print(f"el={el} cells={cells}: tuned {t}", flush=True)
    assert t["candidates"] > 0 and t["sweep_ms"] > 0
    after = (g.positions(), g.last_positions(), g.velocities())
    for a, b in zip(before, after):
        assert np.array_equal(a, b)          # tuning put the state back
    cur = g.get_tile_ex()
    assert {k: cur[k] for k in ("lanes", "qpl", "mode")} == {k: t[k] for k in ("lanes", "qpl", "mode")}
    for _ in range(2):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    print(f"el={el}: 3 steps with the tuned layout, err {err:.3e}", flush=True)
    assert err <= 3 * TOL1
    # a second scene of the same size bucket takes the cached layout
    h, _, _, _ = make_pair(engine, el, cells, young=1e6, poisson=0.45)
    t2 = h.autotune()
    assert t2["candidates"] == 0
    assert {k: t2[k] for k in ("lanes", "qpl", "mode", "inc_per_block")} == \
        {k: t[k] for k in ("lanes", "qpl", "mode", "inc_per_block")}
    recs = [r for r in oe.autotune_cache() if r["element"] == el and r["form"] == "scene"]
    assert len(recs) == 1 and recs[0]["candidates"] == t["candidates"]


def test_explicit_scene_autotune_includes_the_element_pass(engine):
    """The explicit force pass tunes over the gather tiles and the
    element-centric pass (mode 2).  On small beams one gather launch can beat
    the element pass's two launches, so either may win; the tuned layout is the
    scene's and the state is put back."""
    oe.autotune_clear()
    kw = dict(solver=2, material=2, young=1e6, poisson=0.45, noise=1e-3, seed=2)
    g, o, _, _ = make_pair(engine, 9, (8, 2, 2), **kw)
    g.step(1e-3, 10)
    o.step(1e-3, 10)
    before = (g.positions().copy(), g.velocities().copy())
    t = g.autotune()
    print(f"explicit Q2: tuned {t}", flush=True)
    assert t["candidates"] > 1 and t["mode"] in (oe.PASS_GATHER, oe.PASS_TASKS)
    assert g.get_tile_ex()["mode"] == t["mode"]
    for a, b in zip(before, (g.positions(), g.velocities())):
        assert np.array_equal(a, b)
    g.step(1e-3, 10)
    o.step(1e-3, 10)
    assert max_err(g.positions(), o.positions()) <= TOL1


def test_engine_autotune_on_create(engine):
    oe.autotune_clear()
    eng = oe.Engine(0)
    eng.set_autotune(True)
    s = oe.make_scene(eng, oe.TETRA10, (6, 2, 2), young=5e5, poisson=0.49)
    recs = oe.autotune_cache()
    assert len(recs) == 1 and recs[0]["element"] == oe.TETRA10
    cur = s.get_tile_ex()
    assert cur["mode"] == recs[0]["mode"] and cur["lanes"] == recs[0]["lanes"]
    s.step(0.005, 1)
    assert np.isfinite(s.positions()).all()
    s.close()
    eng.close()


def test_batch_autotune_keeps_state_and_matches_alone(engine):
    oe.autotune_clear()
    specs = [(oe.TETRA10, (8, 2, 2), 5e5, 0.49), (oe.HEXA27, (4, 2, 2), 1e6, 0.45),
             (oe.TETRA10, (5, 3, 2), 2e6, 0.3), (oe.HEXA27, (6, 2, 2), 3e5, 0.49)]
    batched = [oe.make_scene(engine, el, c, young=E, poisson=nu) for el, c, E, nu in specs]
    alone = [oe.make_scene(engine, el, c, young=E, poisson=nu) for el, c, E, nu in specs]
    b = oe.Batch(engine, batched)
    b.step(0.005, 1)
    for s in alone:
        s.step(0.005, 1)
    before = [s.positions().copy() for s in batched]
    b.autotune()
    for s, p in zip(batched, before):
        assert np.array_equal(s.positions(), p)
    recs = {r["element"]: r for r in oe.autotune_cache() if r["form"] == "batch"}
    assert set(recs) == {oe.TETRA10, oe.HEXA27}
    for el, r in recs.items():
        cur = b.get_tile_ex(el)
        print(f"batch element {el}: tuned {cur}, {r['ms']:.3f} ms per step over {r['candidates']} layouts",
              flush=True)
        assert cur["mode"] == r["mode"] and cur["lanes"] == r["lanes"] and cur["qpl"] == r["qpl"]
    for _ in range(2):
        b.step(0.005, 1)
        for s in alone:
            s.step(0.005, 1)
    for a, s in zip(batched, alone):
        assert max_err(a.positions(), s.positions()) <= 1e-6
    b.close()
