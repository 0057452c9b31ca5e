"""Writes tests/golden/json_records.npz: the reference's 14 DataRecorder result
files (assets/result/vtk/VBD Small Steps/*.json and
assets/scripts/thesis/XPBD/Torsion_*.json), stored byte for byte, so that the
record writer's output can be compared with what the reference wrote.

Run in a container that has the reference at /root/reference:
    python tests/golden/make_json_fixtures.py
"""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/assets")
DIRS = ["result/vtk/VBD Small Steps", "scripts/thesis/XPBD"]
OUT = Path(__file__).resolve().parent / "json_records.npz"


def main():
    data = {}
    for d in DIRS:
        for f in sorted((SRC / d).glob("*.json")):
            data[f.name] = np.frombuffer(f.read_bytes(), np.uint8)
    np.savez_compressed(OUT, **data)
    print(OUT, {k: v.size for k, v in data.items()})


if __name__ == "__main__":
    main()
