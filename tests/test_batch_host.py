"""CPU checks of the batched-ensemble boundary: the ctypes mirror of isf_member_desc has
the C layout of include/isf.h, bad batches are rejected before any device call, and the
ensemble draw follows its stated distribution."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FIELDS = ("res", "box", "hbar", "dt", "kvec", "omega", "cen", "len", "rad", "n_spawn", "seed", "phase_mode")


def test_member_desc_layout_matches_header(tmp_path):
    from isf_amd import MemberDesc
    src = tmp_path / "layout.c"
    body = "".join(f'    printf("%zu ", offsetof(isf_member_desc, {f}));\n' for f in FIELDS)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "isf.h"\nint main(void) {\n' + body +
                   '    printf("%zu\\n", sizeof(isf_member_desc));\n    return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    c_vals = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    py_vals = [getattr(MemberDesc, f).offset for f in FIELDS] + [C.sizeof(MemberDesc)]
    assert c_vals == py_vals


def test_batch_create_rejects_bad_members_without_gpu():
    import isf_amd
    lib = isf_amd.load()
    h = C.c_void_p()
    assert lib.isf_batch_create(C.byref(h), 0, None) == -1
    arr = (isf_amd.MemberDesc * 2)()
    assert lib.isf_batch_create(C.byref(h), 0, arr) == -1
    arr[0].res = (C.c_int * 3)(60, 32, 32)
    arr[0].box = (C.c_int * 3)(4, 2, 2)
    arr[0].hbar, arr[0].dt = 0.1, 0.02
    rc = lib.isf_batch_create(C.byref(h), 2, arr)
    assert rc == -1 and b"member 0" in lib.isf_last_error()
    arr[0].res = (C.c_int * 3)(64, 256, 128)  # supported by isf_create, too large a plane for a batch
    rc = lib.isf_batch_create(C.byref(h), 2, arr)
    assert rc == -1 and b"up to 128" in lib.isf_last_error()
    assert lib.isf_batch_step(None, 1) == -4
    assert lib.isf_batch_destroy(None) == -4


def test_ensemble_draw_distribution():
    from isf_amd.ensemble import ENSEMBLE_GRIDS, draw_jets
    ms = draw_jets(count=400, seed=3)
    assert {m["res"] for m in ms} == set(ENSEMBLE_GRIDS)
    for m in ms:
        assert m["box"] == tuple(v // 16 for v in m["res"])
        assert 0.05 <= m["hbar"] <= 0.2
        v = np.asarray(m["velocity"], np.float64)
        speed = np.linalg.norm(v)
        assert 0.5 - 1e-6 <= speed <= 1.5 + 1e-6
        assert math.degrees(math.acos(min(1.0, v[0] / speed))) <= 20.0 + 1e-4
        k = np.asarray(m["kvec"], np.float32)
        assert np.allclose(k * m["hbar"], v, rtol=1e-6, atol=1e-6)
        assert np.isclose(m["omega"], speed ** 2 / (2 * m["hbar"]), rtol=1e-5)
    # reproducible for a seed, different seeds per member
    assert draw_jets(count=5, seed=9) == draw_jets(count=5, seed=9)
    assert len({m["seed"] for m in ms}) == len(ms)
