"""The allocation-free residual entry point (oe_scene_residual_scratch)
against the allocating one (oe_scene_residual): same kernel and summation
order, so the results must be bit-identical.  Two calls in a row, the second
on a smaller scene, then the larger one again (the static scratch keeps the
last call's contents, so stale values would show up at vertices the smaller
call leaves alone or at fixed vertices)."""
import numpy as np
import pytest

from scenes import fixed_ids

pytestmark = pytest.mark.gpu

DT = 0.005


def _scene(eng, element, cells):
    import octopusengine_amd as oe

    xyz, topo = oe.beam_mesh(element, cells, (1.0, 0.25, 0.25))
    s = oe.Scene(eng, element, xyz, topo, material=oe.STABLE_NEOHOOKE, young=1e5, poisson=0.45, iterations=3)
    s.fix_plane((1e-6, 0.0, 0.0), (-1.0, 0.0, 0.0))
    s.box_limit((-10, -5, -10), (10, 5, 10))
    for _ in range(2):
        s.step(DT, 1)
    return s, xyz


def _check(s, xyz, label):
    ref = s.residual(DT)
    got = s.residual_scratch(DT)
    print(f"{label}: {s.n_vertices} vertices, max |r| {np.abs(ref).max():.3e}, "
          f"max |diff| {np.abs(got - ref).max():.3e}", flush=True)
    assert np.abs(ref).max() > 0.0
    assert np.array_equal(got, ref)
    assert np.all(got[fixed_ids(xyz)] == 0.0)


def test_residual_scratch_equals_allocating_residual():
    import octopusengine_amd as oe

    eng = oe.Engine(0)
    big, big_xyz = _scene(eng, oe.TETRA10, (16, 4, 4))     # 2673 vertices
    small, small_xyz = _scene(eng, oe.HEXA27, (4, 2, 2))   # 225 vertices
    _check(big, big_xyz, "first call (P2 16x4x4)")
    _check(small, small_xyz, "second call, smaller (Q2 4x2x2)")
    big.step(DT, 1)
    _check(big, big_xyz, "third call, larger again (P2 16x4x4)")
    with pytest.raises(oe.OEError):
        big.residual_scratch(0.0)
    small.close()
    big.close()
    eng.close()
