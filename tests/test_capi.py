"""CPU checks of the drop-in boundary: libisf.so loads, exports every entry point that
include/isf.h declares, and rejects bad arguments before touching the GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "isf.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(isf_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("isf_create", "isf_step", "isf_schroedinger_flow", "isf_pressure_project", "isf_staggered_advect"):
        assert s in syms
    assert len(syms) >= 30


def test_library_exports_every_declared_symbol():
    import isf_amd
    lib = isf_amd.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
        assert s in isf_amd.SIGNATURES, s
    assert set(isf_amd.SIGNATURES) == set(declared_symbols())


def test_create_rejects_bad_sizes_without_gpu():
    import isf_amd
    lib = isf_amd.load()
    h = C.c_void_p()
    box = (C.c_int * 3)(4, 2, 2)
    for dims in [(60, 32, 32), (64, 32, 24), (1024, 32, 32), (64, 1024, 32), (64, 32, 8), (8, 32, 32)]:
        rc = lib.isf_create(C.byref(h), (C.c_int * 3)(*dims), box, 0.1, 0.02)
        assert rc == -1, dims
        assert lib.isf_last_error()
    rc = lib.isf_create(C.byref(h), (C.c_int * 3)(64, 32, 32), (C.c_int * 3)(4, 0, 2), 0.1, 0.02)
    assert rc == -1 and b"physical_size" in lib.isf_last_error()
    rc = lib.isf_create(C.byref(h), (C.c_int * 3)(64, 32, 32), box, -0.1, 0.02)
    assert rc == -1


def test_null_handle_is_an_error():
    import isf_amd
    lib = isf_amd.load()
    assert lib.isf_step(None, None, None, 0, 1) == -4
    assert lib.isf_destroy(None) == -4
    assert b"handle" in lib.isf_last_error()
    with pytest.raises(isf_amd.ISFError):
        isf_amd._lib.check(lib.isf_synchronize(None), "sync")


def test_python_mirror_exposes_reference_names():
    import isf_amd
    for name in ("ISF", "schroedinger_flow_", "normalize_psi_", "pressure_project_", "velocity_one_form_",
                 "isjet", "restrict_velocity_", "add_particles_", "staggered_advect_"):
        assert hasattr(isf_amd, name)
    # the MatlabPort names (src/MatlabPort.jl:38-273, examples/matlabport/simple.jl:6-7)
    from isf_amd import matlabport
    for name in ("ISF", "land", "SchroedingerFlow", "Normalize", "VelocityOneForm", "StaggeredSharp", "Div",
                 "PoissonSolve", "GaugeTransform", "PressureProject", "StaggeredVelocity", "StaggeredAdvect",
                 "Particles", "Keep"):
        assert callable(getattr(matlabport, name))
