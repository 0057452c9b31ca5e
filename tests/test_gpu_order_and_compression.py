"""Two GPU paths that the default workloads never take, against the FP64
oracle:

* the seeded colour shuffle (a Fisher-Yates shuffle per sweep driven by one
  SplitMix64 stream; the reference shuffles every iteration with a
  time-seeded generator, src/GPU/VBD/GPU_VBD_FEM.cu:713-716): the oracle is
  given the same per-sweep orders;
* an element compressed to J ~ 0.05 and ~ 0.01, where det F is re-evaluated
  in FP64 on the GPU.
"""
import numpy as np
import pytest

from conftest import beam_diag
from scenes import make_pair, max_err

DIAG = beam_diag()
TOL1 = 1e-5 * DIAG
M64 = (1 << 64) - 1


def splitmix64(state):
    """SplitMix64 (Steele, Lea, Flood 2014): returns (new state, output)."""
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def shuffled_orders(seed, ncolors, sweeps):
    state, out = seed, []
    for _ in range(sweeps):
        order = list(range(ncolors))
        for i in range(ncolors - 1, 0, -1):
            state, r = splitmix64(state)
            j = r % (i + 1)
            order[i], order[j] = order[j], order[i]
        out.append(order)
    return out


def test_splitmix64_known_answers():
    # the generator's published first outputs from state 0
    state, got = 0, []
    for _ in range(4):
        state, r = splitmix64(state)
        got.append(r)
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


@pytest.mark.gpu
@pytest.mark.parametrize("el,cells", [(7, (8, 2, 2)), (9, (4, 2, 2))])
def test_seeded_shuffle_matches_oracle_with_the_same_orders(engine, el, cells):
    import octopusengine_amd as oe

    seed, steps = 0x5EED1234, 6
    _, o, _, _ = make_pair(engine, el, cells, iterations=1, young=2e5, poisson=0.45)
    xyz, topo = oe.beam_mesh(el, cells, (1.0, 0.25, 0.25))
    g = oe.Scene(engine, el, xyz, topo, material=oe.STABLE_NEOHOOKE, young=2e5, poisson=0.45, iterations=1,
                 shuffle_seed=seed)
    g.fix_plane((1e-6, 0.0, 0.0), (-1.0, 0.0, 0.0))
    g.box_limit((-10, -5, -10), (10, 5, 10))
    _, asc, _, _ = make_pair(engine, el, cells, iterations=1, young=2e5, poisson=0.45)
    orders = shuffled_orders(seed, g.n_colors, steps)
    assert any(od != sorted(od) for od in orders)
    for od in orders:
        o.color_order(od)
        o.step(0.005, 1)
        asc.step(0.005, 1)
        g.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    err_asc = max_err(g.positions(), asc.positions())
    print(f"el={el} shuffled {steps} sweeps: err vs oracle (same orders) {err:.3e}, "
          f"vs oracle (ascending) {err_asc:.3e}", flush=True)
    assert err <= TOL1
    assert err_asc >= 100.0 * max(err, 1e-9)   # the order really changes the result


@pytest.mark.gpu
@pytest.mark.parametrize("ratio", [0.05, 0.01])
def test_compressed_element_fp64_det_path(engine, ratio):
    """One P1 tet in the middle of the 16x4x4 beam scaled about its centroid
    to volume ratio `ratio` (its neighbours deform with it), then one VBD
    step on the GPU and in the oracle."""
    g, o, rest, topo = make_pair(engine, 3, (16, 4, 4), young=1e5, poisson=0.45)
    t = np.asarray(topo).reshape(-1, 4)
    cent = rest[t].mean(axis=1)
    e = int(np.argmin(np.linalg.norm(cent - np.array([0.5, 0.125, 0.125]), axis=1)))
    p0 = rest.astype(np.float64).copy()
    c = p0[t[e]].mean(axis=0)
    p0[t[e]] = c + (p0[t[e]] - c) * ratio ** (1.0 / 3.0)
    p0 = p0.astype(np.float32)
    g.set_state(p=p0, last_p=p0)
    o.set("p", p0.astype(np.float64))
    o.set("last_p", p0.astype(np.float64))
    vol = o.element_diagnostics()["volume"]
    rest_vol = float(np.abs(np.linalg.det(np.stack([rest[t[e, k]] - rest[t[e, 0]] for k in (1, 2, 3)], 1)))) / 6
    j = vol[e] / rest_vol
    print(f"compressed element {e}: J = {j:.4f}", flush=True)
    assert abs(j - ratio) <= 0.01 * ratio
    g.step(0.005, 1)
    o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    moved = max_err(o.positions(), p0.astype(np.float64))
    print(f"J={ratio}: 1-step err {err:.3e} (tol {TOL1:.3e}), largest motion {moved:.3e}", flush=True)
    assert moved > 1e-4
    assert err <= TOL1
