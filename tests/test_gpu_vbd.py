"""VBD on the GPU vs the FP64 oracle (same scene, same colour order).

Tolerance (from the path's accuracy target): max vertex-position error
<= 1e-5 of the bounding-box diagonal after 1 step, <= 1e-3 after 100 steps.
"""
import numpy as np
import pytest

from conftest import beam_diag
from scenes import make_pair, max_err

pytestmark = pytest.mark.gpu

DIAG = beam_diag()
TOL1 = 1e-5 * DIAG
TOL100 = 1e-3 * DIAG

# (element, cells) small enough for the oracle to run in seconds
CASES = [
    (3, (16, 4, 4)),    # P1, the 425-vertex beam
    (7, (8, 2, 2)),     # P2
    (6, (8, 2, 2)),     # Q1
    (9, (4, 2, 2)),     # Q2
    (8, (4, 1, 1)),     # P3
]


def _ids(case):
    return f"el{case[0]}-{case[1][0]}x{case[1][1]}x{case[1][2]}"


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_vbd_one_step(engine, case):
    el, cells = case
    g, o, rest, _ = make_pair(engine, el, cells, material=3, young=5e5, poisson=0.49)
    g.step(0.005, 1)
    o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    print(f"VBD el={el} cells={cells} 1-step err={err:.3e} (tol {TOL1:.3e})", flush=True)
    assert err <= TOL1
    # fixed vertices snapped back and masked
    ids = np.nonzero(rest[:, 0] < 1e-6)[0]
    assert np.all(g.masks()[ids] == 0)
    assert max_err(g.positions()[ids], rest[ids]) <= 1e-6


@pytest.mark.parametrize("case", CASES[:3], ids=[_ids(c) for c in CASES[:3]])
def test_vbd_many_steps(engine, case):
    el, cells = case
    g, o, rest, _ = make_pair(engine, el, cells, material=3, young=5e5, poisson=0.49)
    steps = 100 if el in (3, 6) else 40
    for _ in range(steps):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    disp = float(np.abs(o.positions() - rest).max())
    print(f"VBD el={el} {steps} steps err={err:.3e} tip displacement={disp:.3e}", flush=True)
    assert disp > 1e-3          # the beam really moved
    assert err <= TOL100


def test_vbd_neohooke_and_damping(engine):
    g, o, _, _ = make_pair(engine, 7, (6, 2, 2), material=2, young=1e6, poisson=0.45, damping=1e-3)
    for _ in range(5):
        g.step(0.005, 2)
        o.step(0.005, 2)
    err = max_err(g.positions(), o.positions())
    print(f"VBD NeoHooke damped err={err:.3e}", flush=True)
    assert err <= TOL1 * 5


def test_vbd_tiles_agree(engine):
    """Every gather tile shape computes the same step (up to FP32 summation order)."""
    ref = None
    for tile in [(4, 1), (8, 1), (16, 1), (64, 1), (4, 2), (8, 4)]:
        g, o, _, _ = make_pair(engine, 7, (6, 2, 2), tile=tile)
        g.step(0.005, 1)
        p = g.positions()
        if ref is None:
            o.step(0.005, 1)
            ref = o.positions()
        err = max_err(p, ref)
        print(f"tile {tile} err {err:.3e}", flush=True)
        assert err <= TOL1


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_vbd_cached_gradients(engine, case):
    """The sweep that keeps Ju(e, q) between colours and updates it with each
    vertex move computes the step of the sweep that rebuilds it from the
    nodes, and both match the oracle."""
    el, cells = case
    lanes = 16 if el in (3, 7) else 64
    ref = oref = None
    for cache in (False, True):
        g, o, _, _ = make_pair(engine, el, cells, tile=(lanes, 1, cache))
        assert g.get_tile() == (lanes, 1, cache)
        for _ in range(2):
            g.step(0.005, 1)
        p = g.positions()
        if ref is None:
            for _ in range(2):
                o.step(0.005, 1)
            oref = o.positions()
            e_orc = max_err(p, oref)
            ref = p
        else:
            e_orc = max_err(p, oref)
            e_pair = max_err(p, ref)
            print(f"el={el} cached vs rebuilt {e_pair:.3e}", flush=True)
            assert e_pair <= TOL1 / 10
        print(f"el={el} cache={cache} 2-step err={e_orc:.3e}", flush=True)
        assert e_orc <= TOL1


@pytest.mark.parametrize("curve", [0.0, 0.02])
def test_vbd_q2_element_tasks(engine, curve):
    """Q2 colour pass as element tasks (sum-factorised gradients) + vertex
    solves: same steps as the gather, and the oracle's, on straight (per-element
    records) and curved (per-point records) meshes."""
    import octopusengine_amd as oe

    res = {}
    for mode in (oe.PASS_GATHER, oe.PASS_TASKS):
        g, o, _, _ = make_pair(engine, 9, (4, 2, 2), curve=curve, tile=(64, 1, mode))
        assert g.get_tile()[2] == mode
        assert g.affine() == (curve == 0.0)
        for _ in range(2):
            g.step(0.005, 1)
            o.step(0.005, 1)
        res[mode] = g.positions()
        err = max_err(res[mode], o.positions())
        print(f"Q2 curve={curve} mode={mode} 2-step err={err:.3e}", flush=True)
        assert err <= TOL1
    pair = max_err(res[oe.PASS_GATHER], res[oe.PASS_TASKS])
    print(f"Q2 tasks vs gather {pair:.3e}", flush=True)
    assert pair <= TOL1 / 10


@pytest.mark.parametrize("curve", [0.0, 0.02])
def test_vbd_q2_pencil_tasks(engine, curve):
    """Q2 throughput form with pencil tasks (9 lanes x 3 points, blocks of 14,
    21, 28 or 35 incidences, so that a block's last round is full or partly
    empty) on every colour (no latency form): the same steps as the 27-lane
    throughput form (blocks of 32) and the oracle's, on straight (per-element
    records) and curved (per-point records) meshes.  The pencil sizes compute
    every task alike and sum each vertex in the owners order: their steps are
    bit-identical."""
    import octopusengine_amd as oe

    res = {}
    for inc in (32, 35, 28, 21, 14):
        g, o, _, _ = make_pair(engine, 9, (4, 2, 2), curve=curve)
        g.set_tile_ex(64, 1, oe.PASS_TASKS, inc_per_block=inc, latency_blocks=0)
        assert g.get_tile_ex()["inc_per_block"] == inc
        for _ in range(3):
            g.step(0.005, 1)
            o.step(0.005, 1)
        res[inc] = g.positions()
        err = max_err(res[inc], o.positions())
        print(f"Q2 curve={curve} throughput form, {inc} incidences per block: 3-step err={err:.3e}", flush=True)
        assert err <= TOL1
    for inc in (35, 28, 21, 14):
        pair = max_err(res[32], res[inc])
        print(f"Q2 pencil ({inc} per block) vs 27-lane form {pair:.3e}", flush=True)
        assert pair <= TOL1 / 10
    for inc in (35, 21, 14):
        pair = max_err(res[28], res[inc])
        print(f"Q2 pencil {inc} vs 28 per block {pair:.3e}", flush=True)
        assert pair == 0.0


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_vbd_element_tasks(engine, case):
    """Task form of the colour pass for every element type: same step as the
    gather form and the oracle."""
    import octopusengine_amd as oe

    el, cells = case
    g, o, _, _ = make_pair(engine, el, cells, tile=(16, 1, oe.PASS_TASKS))
    r, _, _, _ = make_pair(engine, el, cells, tile=(16, 1, oe.PASS_GATHER))
    for _ in range(2):
        g.step(0.005, 1)
        r.step(0.005, 1)
        o.step(0.005, 1)
    e_o = max_err(g.positions(), o.positions())
    e_r = max_err(g.positions(), r.positions())
    print(f"tasks el={el} err oracle {e_o:.3e} gather {e_r:.3e}", flush=True)
    assert e_o <= TOL1 and e_r <= TOL1 / 10


def test_vbd_q2_element_tasks_many_steps(engine):
    import octopusengine_amd as oe

    g, o, rest, _ = make_pair(engine, 9, (4, 2, 2), tile=(64, 1, oe.PASS_TASKS))
    g.set_graph(True)
    for _ in range(40):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    disp = float(np.abs(o.positions() - rest).max())
    print(f"Q2 tasks 40 steps (graph) err={err:.3e} displacement={disp:.3e}", flush=True)
    assert disp > 1e-3 and err <= TOL100


def test_gradient_cache_tile_checks(engine):
    import octopusengine_amd as oe

    g, _, _, _ = make_pair(engine, 7, (2, 1, 1))
    with pytest.raises(oe.OEError):
        g.set_tile(16, 1, 3)             # no such mode
    with pytest.raises(oe.OEError):
        g.set_tile(16, 2, True)          # the cache takes one quadrature point per lane
    x, _, _, _ = make_pair(engine, 7, (2, 1, 1), solver=oe.SOLVER_EXPLICIT, iterations=1)
    with pytest.raises(oe.OEError):
        x.set_tile(16, 1, True)          # VBD option only
    x.set_tile(16, 1, False)


def test_vbd_cached_gradients_many_steps(engine):
    g, o, rest, _ = make_pair(engine, 9, (4, 2, 2), tile=(64, 1, True))
    for _ in range(40):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    disp = float(np.abs(o.positions() - rest).max())
    print(f"Q2 cached 40 steps err={err:.3e} displacement={disp:.3e}", flush=True)
    assert disp > 1e-3
    assert err <= TOL100


CURVED = [(7, (6, 2, 2)), (8, (3, 1, 1)), (9, (4, 2, 2)), (6, (6, 2, 2))]


@pytest.mark.parametrize("case", CURVED, ids=[_ids(c) for c in CURVED])
def test_vbd_curved_elements(engine, case):
    """Curved elements keep one record per (element, quadrature point); the
    straight beam of the same element type uses per-element records."""
    el, cells = case
    straight, _, _, _ = make_pair(engine, el, cells)
    assert straight.affine()
    for cache in (False, True):
        lanes = 16 if el == 7 else 64
        g, o, _, _ = make_pair(engine, el, cells, curve=0.02, tile=(lanes, 1, cache))
        assert not g.affine()
        g.step(0.005, 1)
        o.step(0.005, 1)
        err = max_err(g.positions(), o.positions())
        print(f"curved el={el} cache={cache} 1-step err={err:.3e}", flush=True)
        assert err <= TOL1


def test_vbd_chebyshev(engine):
    g, o, _, _ = make_pair(engine, 7, (6, 2, 2), rho=0.7, iterations=6)
    for _ in range(3):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    print(f"VBD Chebyshev rho=0.7 3 steps err={err:.3e}", flush=True)
    assert err <= TOL1


@pytest.mark.parametrize("target", [6, 9])
def test_vbd_better_coloration_jacobi(engine, target):
    """Better_Coloration: DSAT reduced to `target` colours (bit-exact with the
    oracle's), each colour solved from frozen positions then copied."""
    g, o, _, _ = make_pair(engine, 7, (6, 2, 2), vbd_version=3, target_color=target)
    assert g.n_colors == target
    assert np.array_equal(g.colors(), o.colors())
    for _ in range(2):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    print(f"VBD Better_Coloration target={target} err={err:.3e}", flush=True)
    assert err <= TOL1


@pytest.mark.parametrize("case", [(3, (8, 2, 2)), (7, (6, 2, 2)), (9, (4, 2, 2)), (8, (3, 1, 1))],
                         ids=lambda c: f"el{c[0]}")
def test_element_diagnostics_and_residual(engine, case):
    """Per-element stress / volume / volume_diff / inverted and the VBD
    residual forces after a few (unconverged) steps, vs the oracle."""
    el, cells = case
    g, o, _, _ = make_pair(engine, el, cells, iterations=3, young=1e5, poisson=0.45)
    for _ in range(3):
        g.step(0.005, 1)
        o.step(0.005, 1)
    dg, do = g.element_diagnostics(), o.element_diagnostics()
    for k in ("stress", "volume", "volume_diff"):
        scale = float(np.abs(do[k]).max())
        err = float(np.abs(dg[k] - do[k]).max())
        print(f"el={el} {k}: max {scale:.3e} err {err:.3e}", flush=True)
        tol = 1e-4 * scale if k == "stress" else 1e-5 * float(np.abs(do["volume"]).max())
        assert err <= tol
    assert np.array_equal(dg["inverted"], do["inverted"].astype(np.float32))
    rg, ro = g.residual(0.005), o.residual(0.005)
    err = float(np.abs(rg - ro).max())
    print(f"el={el} residual max {np.abs(ro).max():.3e} err {err:.3e}", flush=True)
    assert np.abs(ro).max() > 0
    assert err <= 2e-2 * float(np.abs(ro).max())


def test_vbd_graph_replay_matches_eager(engine):
    a, _, _, _ = make_pair(engine, 9, (4, 2, 2))
    b, _, _, _ = make_pair(engine, 9, (4, 2, 2))
    b.set_graph(True)
    for i in range(4):
        a.step(0.005, 1)
        b.step(0.005, 1)
        engine.synchronize()
        assert max_err(a.positions(), b.positions()) == 0.0, f"replay {i}"


def test_vbd_free_fall_closed_form(engine):
    """lambda = mu ~ 0 (E tiny): every free vertex follows x + g dt^2 per step from rest."""
    g, o, rest, _ = make_pair(engine, 7, (4, 1, 1), young=1e-6, poisson=0.3, fix=False, box=False)
    dt = 0.01
    g.step(dt, 1)
    p = g.positions()
    expect = rest.astype(np.float64).copy()
    expect[:, 1] += -9.81 * dt * dt
    assert max_err(p, expect) <= 1e-6


def _vtk_data(text):
    """{name: array} of the SCALARS / VECTORS blocks of a legacy-VTK file, and
    the block names in order."""
    toks = text.split("\n")
    out, order = {}, []
    n_point = n_cell = 0
    mode = None
    i = 0
    while i < len(toks):
        line = toks[i]
        if line.startswith("POINT_DATA"):
            n_point, mode = int(line.split()[1]), "p"
        elif line.startswith("CELL_DATA"):
            n_cell, mode = int(line.split()[1]), "c"
        elif line.startswith("SCALARS"):
            name = line.split()[1]
            out[name] = np.array(toks[i + 2].split(), np.float32)
            assert out[name].size == (n_point if mode == "p" else n_cell)
            order.append(name)
            i += 2
        elif line.startswith("VECTORS"):
            name = line.split()[1]
            out[name] = np.array(toks[i + 1].split(), np.float32).reshape(-1, 3)
            assert out[name].shape[0] == n_point
            order.append(name)
            i += 1
        i += 1
    return out, order


def test_save_vtk_fem_record_fields(engine, tmp_path):
    """save_vtk writes FEM_VTK_Recorder's fields: weights, u_norm, u,
    smooth_stress (point data); stress, volume, volume_diff, inverted (cell
    data).  Values vs the engine's own getters and the oracle's diagnostics."""
    import octopusengine_amd as oe

    g, o, _, _ = make_pair(engine, oe.TETRA10, (4, 2, 2), iterations=3, young=1e5, poisson=0.45)
    for _ in range(3):
        g.step(0.005, 1)
        o.step(0.005, 1)
    path = tmp_path / "s.vtk"
    g.save_vtk(path)
    d, order = _vtk_data(path.read_text())
    assert order == ["weights", "u_norm", "u", "smooth_stress", "stress", "volume", "volume_diff", "inverted"]
    rest, pos, m = g.init_positions(), g.positions(), g.masses()
    u = pos - rest
    assert np.allclose(d["weights"], m, rtol=1e-5)
    assert np.allclose(d["u"], u, rtol=1e-5, atol=1e-9)
    assert np.allclose(d["u_norm"], np.linalg.norm(u, axis=1), rtol=1e-5, atol=1e-9)
    do = o.element_diagnostics()
    for k in ("stress", "volume", "volume_diff"):
        scale = float(np.abs(do[k]).max())
        tol = 1e-4 * scale if k == "stress" else 1e-5 * float(np.abs(do["volume"]).max())
        assert float(np.abs(d[k] - do[k]).max()) <= tol, k
    assert np.array_equal(d["inverted"], do["inverted"].astype(np.float32))
    # smooth_stress: sum of the incident elements' stress over the vertex mass
    topo = oe.beam_mesh(oe.TETRA10, (4, 2, 2))[1]
    acc = np.zeros(len(rest), np.float64)
    np.add.at(acc, topo.reshape(-1), np.repeat(d["stress"].astype(np.float64), topo.shape[1]))
    assert np.allclose(d["smooth_stress"], acc / m, rtol=1e-4, atol=1e-6 * float(np.abs(acc / m).max()))
