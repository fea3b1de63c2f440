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
