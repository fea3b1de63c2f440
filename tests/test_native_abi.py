"""C-ABI library checks that need no GPU: it loads, exports every entry point the
header declares, and rejects bad configurations before touching the device."""
import ctypes
import subprocess

import pytest

from starhub_amd import native


def test_library_exports_every_declared_symbol():
    path = native.lib_path()
    assert path.exists(), "run `python -m starhub_amd.build` first"
    declared = native.exported_symbols()
    assert len(declared) >= 18
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(str(path))
    for s in declared:
        assert hasattr(lib, s)


def test_status_strings_and_version():
    L = native.lib()
    assert b"gfx950" in L.sh_version()
    for code in range(7):
        assert L.sh_status_string(code)


@pytest.mark.parametrize("field,value", [("knn", 5), ("max_nodes", 5000), ("max_width", 0),
                                         ("pcg_iterations", 0), ("stream_precision", 7), ("spmv_mode", 9)])
def test_create_rejects_bad_config_without_device(field, value):
    from starhub_amd.scene import SolverParams
    cfg = native.make_config(SolverParams(), 64, 48, 128, 8)
    setattr(cfg, field, value)
    h = ctypes.c_void_p()
    rc = native.lib().sh_solver_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == 1  # SH_E_ARG
    assert not h.value


def test_producers_reject_bad_arguments_without_device():
    L = native.lib()
    null = ctypes.c_void_p()
    assert L.sh_skin_surfels(null, 10, null, 10, 5, 1.0, null, null, null, null) == 1
    assert L.sh_build_node_graph(null, 3, 8, 1.0, null, null, null, null, null) == 1


def test_cpp_wrapper_compiles_links_and_reports_errors(tmp_path):
    """The C++ drop-in (include/sh/warp_solver.hpp) builds against the C-ABI
    with the host compiler alone, and a rejected configuration surfaces as
    sh::Error carrying SH_E_ARG -- no GPU needed."""
    root = native.lib_path().parents[2]
    src = tmp_path / "wrap.cpp"
    src.write_text(r'''
#include <sh/warp_solver.hpp>
#include <cstdio>
int main() {
    sh_solver_config c{};
    c.max_width = 64; c.max_height = 48; c.max_nodes = 128; c.knn = 5;  // knn 5: invalid
    c.gn_iterations = 1; c.pcg_iterations = 1;
    sh_camera cam{64, 48, 50.f, 50.f, 32.f, 24.f};
    float pose[16] = {1,0,0,0, 0,1,0,0, 0,0,1,0, 0,0,0,1};
    try {
        sh::OptimizationProcessorWarpSolver proc(c, cam, pose);
        std::puts("constructed");
        return 2;
    } catch (const sh::Error &e) {
        std::printf("%d %s\n", e.status(), e.what());
        return e.status() == SH_E_ARG ? 0 : 3;
    }
}
''')
    exe = tmp_path / "wrap"
    lib = native.lib_path()
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{root / 'include'}", str(src), str(lib),
                    f"-Wl,-rpath,{lib.parent}", "-o", str(exe)], check=True, capture_output=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "invalid argument" in res.stdout
