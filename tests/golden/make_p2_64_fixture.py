"""Writes tests/golden/p2_64x16x16_100steps.npz: the FP64 oracle's state of
the single-scene VBD workload after 100 steps, for the GPU test of the
100-step accuracy bound at full size (the oracle needs about ten minutes for
it, too long to run inside a test).

Scene: P2 beam 64 x 16 x 16 cells of 1.0 x 0.25 x 0.25 m (140,481 vertices),
Stable Neo-Hookean E = 5e5 Pa, nu = 0.49, density 1000, FemShape masses,
x = 0 face fixed, box limit (-10,-5,-10)..(10,5,10), gravity -9.81 y,
dt = 0.005 s, 10 VBD iterations, ascending colour order.

Kept: the displacement (float32) after 100 steps of every vertex whose id is
a multiple of 4 and of every vertex on the free end (x = 1), with their ids,
plus the displacement after 1 step of the same vertices.

Run (CPU only, from the repository root):
    python tests/golden/make_p2_64_fixture.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

OUT = Path(__file__).resolve().parent / "p2_64x16x16_100steps.npz"


def main():
    import octopusengine_amd as oe
    import pyoracle as O

    xyz, topo = oe.beam_mesh(oe.TETRA10, (64, 16, 16), (1.0, 0.25, 0.25))
    o = O.Scene(O.TETRA10, xyz, topo, material=O.STABLE_NEOHOOKE, young=5e5, poisson=0.49, iterations=10)
    o.fix(np.nonzero(xyz[:, 0] < 1e-6)[0])
    o.box((-10, -5, -10), (10, 5, 10))
    ids = np.union1d(np.arange(0, xyz.shape[0], 4), np.nonzero(xyz[:, 0] > 1.0 - 1e-6)[0]).astype(np.int32)
    u1 = None
    for k in range(100):
        o.step(0.005, 1)
        if k == 0:
            u1 = (o.positions()[ids] - xyz[ids]).astype(np.float32)
        print(f"step {k + 1}", flush=True)
    u100 = (o.positions()[ids] - xyz[ids]).astype(np.float32)
    np.savez_compressed(OUT, ids=ids, u1=u1, u100=u100, steps=np.int32(100))
    print(OUT, ids.size, float(np.abs(u100).max()))


if __name__ == "__main__":
    main()
