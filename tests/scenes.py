"""Matched engine / oracle scenes for the parity tests."""
from __future__ import annotations

import numpy as np

import pyoracle as O

# x = 0 face of the beam: Plane(o = (1e-6, 0, 0), n = -x) selects x < 1e-6
FIX_O = (1e-6, 0.0, 0.0)
FIX_N = (-1.0, 0.0, 0.0)
BOX_MIN = (-10.0, -5.0, -10.0)
BOX_MAX = (10.0, 5.0, 10.0)


def fixed_ids(xyz):
    return np.nonzero(xyz[:, 0] < 1e-6)[0].astype(np.int32)


def curve_rest(xyz, amplitude):
    """Smooth non-affine map of the rest shape that keeps x (and so the fixed
    x = 0 face): higher-order nodes leave the straight edges and the elements
    become curved."""
    x = xyz[:, 0].astype(np.float64)
    out = xyz.astype(np.float64).copy()
    out[:, 1] += amplitude * np.sin(3.0 * x) * (1.0 + 4.0 * xyz[:, 2])
    out[:, 2] += amplitude * np.cos(2.0 * x) * xyz[:, 1] * 4.0
    return out.astype(np.float32)


def make_pair(engine, element, cells, size=(1.0, 0.25, 0.25), *, solver=0, material=3, young=5e5, poisson=0.49,
              density=1000.0, iterations=10, damping=0.0, global_damping=0.0, xpbd_version=2, femshape=True,
              noise=0.0, seed=0, box=True, fix=True, tile=None, curve=0.0, rho=0.0, vbd_version=2,
              target_color=-1):
    import octopusengine_amd as oe

    xyz, topo = oe.beam_mesh(element, cells, size)
    rest = curve_rest(xyz, curve) if curve else xyz.copy()
    g = oe.Scene(engine, element, rest, topo, material=material, young=young, poisson=poisson, density=density,
                 distribution=oe.MASS_FEMSHAPE if femshape else oe.MASS_UNIFORM, solver=solver,
                 iterations=iterations, damping=damping, global_damping=global_damping, xpbd_version=xpbd_version,
                 rho=rho, vbd_version=vbd_version, target_color=target_color)
    o = O.Scene(element, rest, topo, material=material, young=young, poisson=poisson, density=density,
                femshape=femshape, solver=solver, iterations=iterations, damping=damping,
                global_damping=global_damping, xpbd_version=xpbd_version)
    if solver == 0 and (rho > 0.0 or vbd_version == 3):
        o.vbd_options(rho=rho, jacobi=vbd_version == 3, target_color=target_color if vbd_version == 3 else -1)
    if fix:
        n = g.fix_plane(FIX_O, FIX_N)
        ids = fixed_ids(rest)
        assert n == ids.size
        o.fix(ids)
    if box:
        g.box_limit(BOX_MIN, BOX_MAX)
        o.box(BOX_MIN, BOX_MAX)
    if noise > 0.0:
        rng = np.random.default_rng(seed)
        p0 = (rest + rng.uniform(-noise, noise, rest.shape)).astype(np.float32)
        g.set_state(p=p0, last_p=p0)
        o.set("p", p0.astype(np.float64))
        o.set("last_p", p0.astype(np.float64))
    if tile is not None:
        g.set_tile(*tile)
    return g, o, rest, topo


def max_err(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max())
