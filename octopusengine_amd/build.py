"""Build the engine's shared library (liboe.so) for gfx950 with hipcc.

Objects go to octopusengine_amd/_build/, the library to
octopusengine_amd/liboe.so (in-tree, so it travels with the source tree).
Run as ``python octopusengine_amd/build.py`` (a path, not ``-m``: importing the
package loads the library being rebuilt) or call :func:`build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = PKG / "liboe.so"
ARCH = os.environ.get("OE_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

COMMON = ["-O3", "-std=c++17", "-fPIC"]
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                      "-fno-gpu-rdc"]
# Host-only geometry / graph code: no FMA contraction, so float expressions are
# rounded term by term.
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off"]

HOST_SOURCES = ["oe_shapes.cpp", "oe_mesh.cpp", "oe_color.cpp", "oe_fem.cpp", "oe_vtk.cpp", "oe_loader.cpp",
                "oe_record.cpp"]
HIP_SOURCES = ["oe_gpu.cpp", "oe_scene.cpp", "oe_batch.cpp", "oe_server.cpp", "oe_capi.cpp", "oe_elementwise.hip", "oe_xpbd.hip", "oe_diag.hip", "oe_q2.hip", "oe_q2p.hip", "oe_scratch.hip", "oe_tune.cpp"]
GATHER_VARIANTS = [(4, 1), (8, 8), (10, 4), (20, 11), (27, 27)]
EXTRA_HIP_SOURCES: list[str] = []


def _jobs() -> list[tuple[list[str], Path]]:
    jobs = []
    for src in HOST_SOURCES:
        obj = OUT_DIR / (Path(src).stem + ".o")
        jobs.append(([CXX, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], obj))
    for src in HIP_SOURCES + EXTRA_HIP_SOURCES:
        obj = OUT_DIR / (Path(src).stem + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        jobs.append(([HIPCC, *HIP_FLAGS, *lang, "-c", str(CSRC / src), "-o", str(obj)], obj))
    for s, q in GATHER_VARIANTS:
        obj = OUT_DIR / f"oe_gather_{s}_{q}.o"
        defs = [f"-DOE_GATHER_S={s}", f"-DOE_GATHER_Q={q}"]
        jobs.append(([HIPCC, *HIP_FLAGS, *defs, "-c", str(CSRC / "oe_gather.hip"), "-o", str(obj)], obj))
    return jobs


def _run(cmd: list[str]) -> tuple[int, str]:
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return p.returncode, p.stdout


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source and link liboe.so; raises RuntimeError on failure."""
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    work = _jobs()
    n = jobs or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=n) as ex:
        results = list(ex.map(lambda j: (_run(j[0]), j[0]), work))
    errors = []
    for (rc, out), cmd in results:
        # hipcc's host pass does not know the device-only target feature; that
        # warning is expected and filtered.
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if rc != 0:
            errors.append(" ".join(cmd) + "\n" + "\n".join(lines))
        elif verbose and lines:
            print("\n".join(lines))
    if errors:
        raise RuntimeError("build failed:\n" + "\n\n".join(errors))
    objs = [str(o) for _, o in work]
    tmp = LIB.with_suffix(".so.tmp")
    rc, out = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs])
    if rc != 0:
        raise RuntimeError("link failed:\n" + out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
