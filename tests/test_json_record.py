"""JSON experiment record (the reference's DataRecorder) and the flexion /
rotation error measures it holds.  CPU only.

Golden files: tests/golden/json_records.npz holds the reference's own 14
record files (written by tests/golden/make_json_fixtures.py).  Each is rebuilt
entry by entry from its parsed values and must come out byte for byte; for the
"VBD Small Steps" beams the mesh entry is taken from the engine's own beam
generator, which pins its element and vertex counts too.  The error measures
are pinned by closed forms: a rigid translation for the flexion error, a rigid
rotation about the beam axis for the rotation error."""
import json
from pathlib import Path

import numpy as np
import pytest

import octopusengine_amd as oe

GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "json_records.npz")
NAMES = sorted(GOLDEN.files)
ELEMENTS = {"Tetra": oe.TETRA, "Hexa": oe.HEXA, "Tetra10": oe.TETRA10, "Tetra20": oe.TETRA20, "Hexa27": oe.HEXA27}


def test_fixture_count():
    assert len(NAMES) == 14


@pytest.mark.parametrize("name", NAMES)
def test_record_bytes(tmp_path, name):
    text = GOLDEN[name].tobytes().decode()
    doc = json.loads(text)
    rec = oe.Recorder(tmp_path / "r.json")
    for key, val in doc.items():
        if key == "Time":
            rec.time(val["frame"], val["time"], val["dt"], val["fixed_Time"], val["fixed_dt"])
        elif key == "mesh":
            counts = {ELEMENTS[k]: v for k, v in val.items() if k != "vertices"}
            if "Hexa_16_4_4_4x1x1" in name:
                xyz, topo = oe.beam_mesh(oe.HEXA, (16, 4, 4), (4.0, 1.0, 1.0))
                assert counts == {oe.HEXA: topo.shape[0]} and val["vertices"] == xyz.shape[0]
            rec.mesh(counts, val["vertices"])
        elif key == "rotation_error":
            rec.rotation(val["rot_max"], val["beam_lenght"], val["distance"], val["angles"])
        elif key == "flexion_error":
            rec.flexion(val["error"], val["p"], val["target"])
        else:
            assert isinstance(val, str), key
            rec.string(key, val)
    assert rec.text() == text
    rec.save()
    assert (tmp_path / "r.json").read_bytes() == GOLDEN[name].tobytes()


def test_material_entries():
    rec = oe.Recorder("/dev/null")
    rec.material(oe.STABLE_NEOHOOKE, 0.49, 5e5, 1, 1000.0)
    rec.material(oe.NEOHOOKE, 0.45, 1e6, 20, 1000.0, iterations=1)
    assert rec.text() == (
        '{"XPBD_FEM_Dynamic" : {"material" : "Stable Neo-Hooke","poisson" : 0.49,"young" : 500000,"sub_step" : 1,'
        '"density" : 1000},"XPBD_FEM_Dynamic" : {"material" : "Neo-Hooke","poisson" : 0.45,"young" : 1e+06,'
        '"iteration" : 1,"sub_itetrations" : 20,"density" : 1000}}')


def test_flexion_translation():
    xyz, _ = oe.beam_mesh(oe.TETRA10, (8, 2, 2), (4.0, 1.0, 1.0))
    shift = np.array([0.0, -0.25, 0.125], np.float32)
    follow = np.array([4.0, 0.5, 0.5], np.float32)
    target = np.array([4.0, 0.0, 0.5], np.float32)
    err, p, v = oe.flexion_error(xyz, xyz + shift, follow, target)
    assert v >= 0 and np.allclose(xyz[v], follow, atol=1e-3)
    assert np.allclose(p, xyz[v] + shift)
    assert err == pytest.approx(float(np.sum((target - p) ** 2)), rel=1e-6)
    assert oe.flexion_error(xyz, xyz, [9.0, 9.0, 9.0], target)[2] == -1


@pytest.mark.parametrize("el", [oe.TETRA, oe.HEXA, oe.TETRA10, oe.HEXA27])
@pytest.mark.parametrize("theta", [30.0, 135.0])
def test_rotation_rigid(el, theta):
    """Rotating the beam rigidly about its axis (y, z) = (0.5, 0.5) by theta:
    every sample reads theta, and the distances are the edge samples' x / L."""
    L = 4.0
    xyz, topo = oe.beam_mesh(el, (8, 2, 2), (L, 1.0, 1.0))
    t = np.radians(theta)
    c, s = np.cos(t), np.sin(t)
    y, z = xyz[:, 1].astype(np.float64) - 0.5, xyz[:, 2].astype(np.float64) - 0.5
    pos = xyz.astype(np.float64).copy()
    pos[:, 1] = 0.5 + c * y - s * z
    pos[:, 2] = 0.5 + s * y + c * z
    dist, ang = oe.rotation_error(el, topo, xyz, pos.astype(np.float32), 3, L)
    assert np.allclose(ang, theta, atol=2e-3)
    # the edge y = z = 0 is cut into 8 cells: samples at x = k L / 16 (3 per cell edge, shared ends merged)
    assert np.allclose(dist, np.arange(17) / 16.0, atol=1e-6)
    assert np.all(np.diff(dist) > 0)
