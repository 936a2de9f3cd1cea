"""C ABI checks that need no GPU: the library loads, exports every function
include/collision_amd.h declares, reports a missing device as a status (not a
crash), and the ctypes mirrors match the header's struct layouts."""
import ctypes as C
import os
import subprocess

import pytest

from collision_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_default_params_are_the_reference_defaults():
    p = _lib.Params()
    _lib.lib().cs_default_params(C.byref(p))
    # dcollid.cpp:28-35 and collid.h:17-19
    assert (p.dt, p.thickness, p.round_eps, p.spring_k, p.mass, p.friction, p.restitution) == \
        (1e-3, 1e-4, 1e-6, 1000.0, 0.01, 0.02, 0.0)
    assert p.max_ccd_passes == 5 and p.max_strain_sweeps == 10
    # library switches: incremental passes on, every per-pass time recorded
    assert p.force_full_passes == 0 and p.pass_timing == 2


def test_create_without_device_is_a_status_not_a_crash():
    L = _lib.lib()
    h = C.c_void_p()
    p = _lib.Params()
    L.cs_default_params(C.byref(p))
    rc = L.cs_create(C.byref(h), C.byref(p))
    if rc == _lib.CS_OK:
        L.cs_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == _lib.CS_ERR_NO_DEVICE
    assert b"device" in L.cs_last_error(None)
    assert L.cs_resolve(None, None) == _lib.CS_ERR_ARG


def test_struct_layouts_match_header(tmp_path):
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "abi_layout.c"), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    vals = dict(ln.rsplit(" ", 1) for ln in lines if ln.strip())
    assert int(vals["cs_params"]) == C.sizeof(_lib.Params)
    assert int(vals["cs_mesh_view"]) == C.sizeof(_lib.MeshView)
    assert int(vals["cs_pass_info"]) == C.sizeof(_lib.PassInfo)
    assert int(vals["cs_step_report"]) == C.sizeof(_lib.StepReport)
    for key, val in vals.items():
        if "." in key:
            struct, field = key.split(".")
            cls = {"cs_params": _lib.Params, "cs_mesh_view": _lib.MeshView,
                   "cs_pass_info": _lib.PassInfo, "cs_step_report": _lib.StepReport}[struct]
            assert getattr(cls, field).offset == int(val), key


def test_cpp_host_api_compiles():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "cloth_on_box.cpp")], check=True)
