"""isf_norm2_noalloc: the FP64 norm of isf_norm2 computed with no device allocation (its
partial sums in a fixed per-GPU device array, the sum written to caller memory), called
through the C ABI and compared with isf_norm2 on two calls in a row, the second with a
smaller input; size limits are rejected before any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_norm2_noalloc_matches_norm2():
    from isf_amd import ISF, DeviceArray, load
    lib = load()
    out = DeviceArray((1,), np.float64)
    rng = np.random.default_rng(42)
    for dims in ((64, 32, 32), (32, 16, 16)):  # second call: smaller input
        isf = ISF((4, 2, 2), dims, 0.1, 1 / 48)
        nz, ny, nx = dims[2], dims[1], dims[0]
        host = rng.standard_normal((nz, ny, nx, 4)).astype(np.float32)
        psi = isf.psi_array(host)
        ref = isf.norm2(psi)  # the existing entry point (handle partials)
        rc = lib.isf_norm2_noalloc(psi.ptr, nx * ny * nz, out.ptr, None)
        assert rc == 0, lib.isf_last_error().decode()
        got = float(out.to_host()[0])
        exact = float(np.sum(host.astype(np.float64) ** 2))
        print(f"  {dims}: no-alloc {got:.17g}, isf_norm2 {ref:.17g}, numpy {exact:.17g}")
        assert got == ref  # same kernels, same partition, same order
        assert abs(got - exact) <= 1e-12 * exact
        psi.free()
        isf.close()


def test_norm2_noalloc_rejects_bad_arguments():
    from isf_amd import load
    lib = load()
    dummy = C.c_void_p(16)
    assert lib.isf_norm2_noalloc(None, 10, dummy, None) == -1
    assert lib.isf_norm2_noalloc(dummy, 10, None, None) == -1
    assert lib.isf_norm2_noalloc(dummy, 0, dummy, None) == -1
    assert lib.isf_norm2_noalloc(dummy, (1 << 30) + 1, dummy, None) == -1
    assert b"outside" in lib.isf_last_error()
