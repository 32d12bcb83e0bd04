"""The edge angle of the velocity one-form (Julia's angle, src/parallel.jl:108) is evaluated
on the GPU by `angle_of` in csrc/isf_device.hpp: octant reduction to a = min/max in [0, 1]
and atan(a) = a P(a^2) with a degree-8 polynomial.  This CPU test reads the coefficients
out of that source and checks the evaluation, emulated in FP32 with fused multiply-adds,
against numpy's FP64 arctan2: the polynomial's stated relative error bound on [0, 1] and
the absolute error of the full angle over all octants, signed zeros and the axes."""
import pathlib
import re

import numpy as np

SRC = pathlib.Path(__file__).resolve().parents[1] / "schroedingerssmoke.jl_amd" / "csrc" / "isf_device.hpp"


def _coefficients():
    text = SRC.read_text()
    body = text[text.index("float angle_of(float re, float im)"):]
    body = body[:body.index("return copysignf")]
    lead = re.search(r"float p = ([-+0-9.eE]+)f;", body)
    steps = re.findall(r"p = fmaf\(p, s, ([-+0-9.eE]+)f\);", body)
    assert lead and len(steps) == 8, "angle_of's polynomial is no longer a degree-8 Horner chain"
    return [np.float32(lead.group(1))] + [np.float32(c) for c in steps]


def _fma32(a, b, c):
    # the FP64 product of two FP32 values is exact, so one rounding of a*b + c to FP32
    # is the fused multiply-add except for rare double-rounding ties
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def _atan01(a, coef):
    a = a.astype(np.float32)
    s = (a * a).astype(np.float32)
    p = np.full_like(a, coef[0])
    for c in coef[1:]:
        p = _fma32(p, s, np.full_like(a, c))
    return (p * a).astype(np.float32)


def _angle(re, im, coef):
    re = re.astype(np.float32)
    im = im.astype(np.float32)
    ax, ay = np.abs(re), np.abs(im)
    mx = np.maximum(np.maximum(ax, ay), np.float32(1.17549435e-38))
    mn = np.minimum(ax, ay)
    a = (mn / mx).astype(np.float32)  # v_rcp_f32 is within 1 ulp of 1/mx; the tolerance allows it
    r = _atan01(a, coef)
    r = np.where(ay > ax, np.float32(np.pi / 2) - r, r).astype(np.float32)
    r = np.where(np.signbit(re), np.float32(np.pi) - r, r).astype(np.float32)
    return np.copysign(r, im)


def test_polynomial_relative_error_on_unit_interval():
    coef = _coefficients()
    a = np.linspace(0.0, 1.0, 200001, dtype=np.float64).astype(np.float32)
    got = _atan01(a, coef).astype(np.float64)
    ref = np.arctan(a.astype(np.float64))
    nz = a > 0
    rel = np.abs(got[nz] - ref[nz]) / ref[nz]
    assert rel.max() <= 1.5e-7, rel.max()
    assert got[0] == 0.0


def test_full_angle_against_arctan2():
    coef = _coefficients()
    rng = np.random.default_rng(7)
    th = rng.uniform(-np.pi, np.pi, 100000)
    rad = 10.0 ** rng.uniform(-20, 20, th.size)
    re = (rad * np.cos(th)).astype(np.float32)
    im = (rad * np.sin(th)).astype(np.float32)
    got = _angle(re, im, coef).astype(np.float64)
    ref = np.arctan2(im.astype(np.float64), re.astype(np.float64))
    assert np.abs(got - ref).max() < 6e-7  # a few FP32 ulps of pi
    # axes, signed zeros and the origin: atan2's special values
    re0 = np.array([1, -1, 0, 0, 0, -0.0, 0, -0.0, 3, -3], np.float32)
    im0 = np.array([0, 0, 1, -1, 0, 0, -0.0, -0.0, -0.0, -0.0], np.float32)
    got0 = _angle(re0, im0, coef).astype(np.float64)
    ref0 = np.arctan2(im0.astype(np.float64), re0.astype(np.float64))
    np.testing.assert_allclose(got0, ref0, atol=3e-7)
    assert np.array_equal(np.signbit(got0), np.signbit(ref0))
