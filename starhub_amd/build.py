"""Builds the native libraries in-tree.

* ``starhub_amd/_build/libstarhub_amd.so`` -- the HIP solver library (gfx950),
  exporting the C-ABI of ``include/sh/warp_solver.h``;
* ``starhub_amd/_build/libsh_scene.so`` -- the synthetic scene generator (host C++);
* ``oracle/_build/libsh_oracle.so`` -- the CPU restatement used by the tests and
  the CPU-baseline leg of the benchmark (built by ``oracle/Makefile``).

Every call recompiles.  Run ``python -m starhub_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # packed FP32 (v_pk_*_f32) stays off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unknown-pragmas",
]

HIP_SOURCES = ["sh_kernels.hip", "sh_rows.hip", "sh_tiles.hip", "sh_pcg.hip", "sh_sort.hip", "sh_warp.hip", "sh_measure.hip", "sh_render.hip", "sh_flow.hip", "sh_bpcg.hip",
               "sh_engine.cpp", "sh_server.cpp"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")


def build_hip(jobs: int = 4) -> Path:
    OUT.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = OUT / (path.stem + ".o")
        lang = ["-x", "hip"] if path.suffix == ".cpp" else []
        cmds.append([HIPCC, *HIP_FLAGS, *lang, "-c", str(path), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    lib = OUT / "libstarhub_amd.so"
    tmp = OUT / "libstarhub_amd.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, lib)
    for o in objs:
        o.unlink(missing_ok=True)
    return lib


def build_scene() -> Path:
    OUT.mkdir(exist_ok=True)
    lib = OUT / "libsh_scene.so"
    tmp = OUT / "libsh_scene.so.tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-std=c++17", "-shared",
          str(PKG / "synth" / "scene.cpp"), "-o", str(tmp)])
    os.replace(tmp, lib)
    return lib


def build_oracle() -> Path | None:
    mk = ROOT / "oracle" / "Makefile"
    if not mk.exists():
        return None
    _run(["make", "-s", "-B", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "libsh_oracle.so"


def build_dropin_check() -> Path:
    """tests/cpp/_build/dropin_check: a C++ caller of the drop-in header, run by
    the GPU tests (test infrastructure; host compiler only, links the library)."""
    src = ROOT / "tests" / "cpp" / "dropin_check.cpp"
    out_dir = ROOT / "tests" / "cpp" / "_build"
    out_dir.mkdir(exist_ok=True)
    exe = out_dir / "dropin_check"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}", "-I/opt/rocm/include",
          str(src), str(OUT / "libstarhub_amd.so"), "-L/opt/rocm/lib", "-lamdhip64",
          "-Wl,-rpath,$ORIGIN/../../../starhub_amd/_build:/opt/rocm/lib", "-o", str(exe)])
    return exe


def build_all() -> None:
    build_scene()
    build_oracle()
    build_hip()
    build_dropin_check()


if __name__ == "__main__":
    build_all()
    print("built:", *(p.name for p in sorted(OUT.glob("*.so"))))
