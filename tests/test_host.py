"""CPU-only checks of the host side: the C-ABI library loads and exports every
entry point include/mfx.h declares, and the manta module's argument handling,
vec3 arithmetic and solver bookkeeping follow the reference's wrappers."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "mfx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|void|const char\*)\s+(mfx_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mantaflow_amd import _abi
    names = header_functions()
    assert len(names) >= 45
    L = C.CDLL(_abi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers exactly the header
    assert sorted(_abi.SIGNATURES) == names


def test_oracle_is_not_linked_into_the_library():
    from mantaflow_amd import _abi
    data = open(_abi.LIB_PATH, "rb").read()
    assert b"og_solve_pressure" not in data and b"liboracle" not in data


def test_parse_follows_reference_wrapper_rules():
    from mantaflow_amd.manta import _REQ, _parse
    sig = [("flags", _REQ), ("vel", _REQ), ("order", 1), ("strength", 1.0)]
    a = _parse(sig, ("F", "V"), {"strength": 0.5, "notiming": True})
    assert a == {"flags": "F", "vel": "V", "order": 1, "strength": 0.5}
    with pytest.raises(RuntimeError, match="Argument 'bogus' unknown"):
        _parse(sig, ("F", "V"), {"bogus": 1})
    with pytest.raises(RuntimeError, match="Argument 'vel' is not defined."):
        _parse(sig, ("F",), {})
    with pytest.raises(RuntimeError, match="does not read argument number"):
        _parse(sig, ("F", "V", 1, 2, 3), {})


def test_vec3_and_constants():
    from mantaflow_amd import manta as m
    gs = m.vec3(64, 96, 64)
    c = gs * m.vec3(0.5, 0.1, 0.5)
    assert (c.x, c.y, c.z) == (32.0, 9.6000000000000014, 32.0) or np.allclose(list(c), [32, 9.6, 32])
    assert list(m.vec3(2) + 1) == [3.0, 3.0, 3.0]
    assert (m.FlagFluid, m.FlagObstacle, m.FlagEmpty, m.FlagOutflow) == (1, 2, 4, 16)
    assert (m.PcNone, m.PcMIC, m.PcMGDynamic, m.PcMGStatic) == (0, 1, 2, 3)
    assert m.Solver is m.FluidSolver and m.Vec3 is m.vec3


def test_pressure_defaults_match_reference():
    from mantaflow_amd import _abi
    o = _abi.PressureOpts()
    _abi.lib().mfx_pressure_opts_default(C.byref(o))
    assert abs(o.cgAccuracy - 1e-3) < 1e-9 and abs(o.cgMaxIterFac - 1.5) < 1e-7
    assert o.preconditioner == 1 and o.enforceCompatibility == 0 and o.zeroPressureFixing == 0


def test_errors_surface_as_messages_without_a_device():
    """Without a usable GPU the scene cannot be created; the C ABI reports an
    error string instead of crashing."""
    from mantaflow_amd import _abi
    L = _abi.lib()
    cnt = C.c_int(-1)
    assert L.mfx_device_count(C.byref(cnt)) == 0
    if cnt.value > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = L.mfx_scene_create(16, 16, 16, 3, C.byref(h))
    assert rc == -1 and L.mfx_last_error()
    assert L.mfx_scene_create(16, 16, 2, 2, C.byref(h)) == -1
    assert b"2D solver" in L.mfx_last_error()


def test_serve_engine_host_logic():
    """The serving engine's bookkeeping without a device: option checks,
    requests for unknown sessions refused, poll/drain with nothing
    outstanding, the clock, and shutdown."""
    from mantaflow_amd import _abi
    L = _abi.lib()
    o = _abi.ServeOpts()
    L.mfx_serve_opts_default(C.byref(o))
    assert (o.maxBatch, o.workers, o.hbmReserve, o.memoryLimit) == (64, 1, 2 << 30, 0)
    assert abs(o.solve.cgAccuracy - 1e-4) < 1e-9 and o.solve.preconditioner == 3
    h = C.c_void_p()
    bad = _abi.ServeOpts.from_buffer_copy(o)
    bad.workers = 0
    assert L.mfx_serve_create(C.byref(bad), C.byref(h)) == -1
    assert b"workers" in L.mfx_last_error()
    o.workers = 3
    assert L.mfx_serve_create(C.byref(o), C.byref(h)) == 0
    try:
        assert L.mfx_serve_submit(h, 0, _abi.REQ_SOLVE, 7) == -1
        assert b"unknown session" in L.mfx_last_error()
        assert L.mfx_serve_submit(h, 0, 5, 7) == -1
        assert b"kind" in L.mfx_last_error()
        ss, kk, tt = (C.c_int * 2)(0, 0), (C.c_int * 2)(1, 1), (C.c_longlong * 2)(1, 2)
        assert L.mfx_serve_submit_many(h, 2, ss, kk, tt) == -1  # nothing queued
        buf = (_abi.ServeResult * 4)()
        n = C.c_int(-1)
        assert L.mfx_serve_poll(h, buf, 4, C.byref(n), 0.01) == 0 and n.value == 0
        assert L.mfx_serve_drain(h, 0.01) == 0
        t1, t2 = C.c_double(), C.c_double()
        assert L.mfx_serve_time(h, C.byref(t1)) == 0
        assert L.mfx_serve_time(h, C.byref(t2)) == 0 and 0.0 <= t1.value <= t2.value
        s = _abi.ServeSession()
        assert L.mfx_serve_add_session(h, C.byref(s), None) == -1
        assert b"required" in L.mfx_last_error()
    finally:
        assert L.mfx_serve_destroy(h) == 0


def test_header_compiles_as_c_and_host_program_as_cpp(tmp_path):
    """include/mfx.h is plain C (C99) and the C++ host program of the C-ABI
    tests compiles against it (syntax only, no device needed)."""
    import subprocess
    src = tmp_path / "abi.c"
    names = header_functions()
    src.write_text('#include "mfx.h"\ntypedef void (*fn)(void);\nfn table[] = {%s};\n'
                   'int main(void) { return table[0] == 0; }\n' % ", ".join("(fn)&%s" % n for n in names))
    inc = os.path.join(ROOT, "include")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-pedantic", "-fsyntax-only", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc,
                        os.path.join(ROOT, "tests", "cabi", "plume_host.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
