"""CPU pins of the oracle's MatlabPort particle store (oracle/matlabport.py, restating
src/MatlabPort.jl:178-230 and examples/matlabport/simple.jl:63-70,113-117), and the C
ABI's argument checks for the device store."""
import ctypes as C

import numpy as np
import pytest

from oracle import matlabport as mp


def test_append_grows_length_and_active_and_refuses_the_last_slot():
    p = mp.Particles(10)
    p.append([1.0, 2.0], [3.0, 4.0], [5.0, 6.0])
    p.append([7.0] * 3, [8.0] * 3, [9.0] * 3)
    assert p.length == 5 and p.active == [0, 1, 2, 3, 4]
    assert p.x[:5].tolist() == [1.0, 2.0, 7.0, 7.0, 7.0]
    with pytest.raises(AssertionError):  # length(x) + p.length < length(p.x), src/MatlabPort.jl:187
        p.append([0.0] * 5, [0.0] * 5, [0.0] * 5)
    p.append([0.0] * 4, [0.0] * 4, [0.0] * 4)  # 9 < 10 is allowed
    assert p.length == 9


def test_remove_dead_keeps_order_and_is_strict():
    p = mp.Particles(20)
    xs = [0.5, 0.0, 1.5, 4.0, 3.99, -0.1, 2.0]
    p.append(xs, [1.0] * 7, [1.0] * 7)
    p.remove_dead((4, 2, 2))
    assert p.active == [0, 2, 4, 6]
    p.remove_dead((4, 2, 2))  # idempotent
    assert p.active == [0, 2, 4, 6]


def test_ring_births_and_keep():
    p = mp.Particles(100)
    cen, rad = (0.3, 0.966, 1.066), 0.5
    p.spawn_ring(40, cen, rad, seed=7, step=3)
    r = np.hypot(p.y[:40] - np.float32(cen[1]), p.z[:40] - np.float32(cen[2]))
    assert np.all(p.x[:40] == np.float32(cen[0]))
    assert np.allclose(r, 0.9 * rad, atol=1e-6)
    ang = np.arctan2(p.z[:40] - cen[2], p.y[:40] - cen[1])
    assert np.ptp(ang) > 5.0  # spread over the whole ring
    q = mp.Particles(100)
    q.spawn_ring(40, cen, rad, seed=7, step=3)
    assert np.array_equal(p.y, q.y)  # seeded
    p.keep([5, 1, 3])
    assert p.x.shape == (3,) and p.y[1] == q.y[1]


def test_advect_in_a_constant_field_is_exact():
    p = mp.Particles(10)
    p.append([1.0, 2.0], [0.5, 0.7], [0.5, 0.25])
    n, L = (16, 8, 8), (4, 2, 2)
    t = mp.Torus(n, L, tuple(L[a] / n[a] for a in range(3)))
    vx, vy, vz = np.full(n, 0.3), np.full(n, -0.2), np.full(n, 0.1)
    p.advect(t, vx, vy, vz, 0.5)
    assert np.allclose(p.x[:2], [1.15, 2.15], atol=1e-6)
    assert np.allclose(p.y[:2], [0.4, 0.6], atol=1e-6)
    assert np.allclose(p.z[:2], [0.55, 0.30], atol=1e-6)


def test_device_store_argument_checks_without_gpu():
    import isf_amd
    lib = isf_amd.load()
    h = C.c_void_p()
    assert lib.isf_particles_create(C.byref(h), 0) == -1
    assert lib.isf_particles_create(None, 10) == -1
    assert lib.isf_particles_destroy(None) == -4
    assert lib.isf_particles_remove_dead(None, None) == -4
