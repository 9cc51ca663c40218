"""The FP64 elementary functions of the corner terms (prism_device.hpp).  atan2_f64: its
polynomial coefficients, read from the header, give atan(v) on the reduced range
|v| <= tan(pi/8) within an ulp of numpy's arctan in double arithmetic, and the octant/quadrant
assembly reproduces arctan2 in every octant.  log_pos_f64: its polynomial q matches
(atanh(f)/f - 1)/f^2 = sum f^2k / (2k + 3), in extended precision, to 5e-16 relative over the
reduced range.  The
kernels' results with both are compared with the oracle on the GPU (test_gpu_parity.py,
test_gpu_configs.py)."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "jif3d_amd", "csrc", "prism_device.hpp")


def constants():
    src = open(HDR).read()
    body = src[src.index("double atan2_f64("):]
    body = body[:body.index("\n}\n")]
    first = float(re.search(r"double g = (\S+);", body).group(1))
    rest = [float(v) for v in re.findall(r"g = fma\(g, x, (\S+)\);", body)]
    hi, lo = (float(v) for v in re.search(r"fma\(dm, (\S+), fma\(dm, (\S+), s \* at\)\)", body).groups())
    return [first] + rest, hi, lo


def model(p, q, coef, hi, lo):
    ap, aq = np.abs(p), np.abs(q)
    swap = ap > aq
    num, den = np.minimum(ap, aq), np.maximum(ap, aq)
    big = num > 0.41421356237309503 * den
    n2, d2 = np.where(big, num - den, num), np.where(big, num + den, den)
    v = n2 / d2
    x = v * v
    g = np.full_like(x, coef[0])
    for c in coef[1:]:
        g = g * x + c
    at = v + (v * x) * g
    m = np.where(big, 1.0, 0.0)
    m = np.where(swap, 2.0 - m, m)
    m = np.where(q < 0, 4.0 - m, m)
    s = np.where(swap != (q < 0), -1.0, 1.0)
    res = m * hi + (m * lo + s * at)
    return np.where(p < 0, -res, res)


def test_polynomial_on_the_reduced_range():
    coef, hi, lo = constants()
    assert len(coef) == 11 and abs(hi + lo - np.pi / 4) < 1e-17 and (hi * 2**50).is_integer()
    v = np.linspace(-np.tan(np.pi / 8), np.tan(np.pi / 8), 200001)
    x = v * v
    g = np.full_like(x, coef[0])
    for c in coef[1:]:
        g = g * x + c
    at = v + (v * x) * g
    ref = np.arctan(v)
    nz = ref != 0
    assert np.max(np.abs(at - ref)[nz] / np.spacing(np.abs(ref[nz]))) <= 1.0


def test_every_octant_within_two_ulp():
    coef, hi, lo = constants()
    rng = np.random.default_rng(0)
    p = rng.normal(size=400000) * np.exp(rng.uniform(-20, 20, 400000))
    q = rng.normal(size=400000) * np.exp(rng.uniform(-20, 20, 400000))
    p[:2000] = q[:2000] * np.tan(np.pi / 8) * (1 + rng.uniform(-1e-9, 1e-9, 2000))  # octant edges
    p[2000:4000] = q[2000:4000] * (1 + rng.uniform(-1e-9, 1e-9, 2000))
    f, ref = model(p, q, coef, hi, lo), np.arctan2(p, q)
    assert np.max(np.abs(f - ref) / np.spacing(np.abs(ref))) <= 2.0
    # a wrong octant or quadrant offset would be off by a multiple of pi/4
    assert np.all(np.sign(f) == np.sign(ref))


def log_coefficients():
    src = open(HDR).read()
    body = src[src.index("double log_pos_f64("):]
    body = body[:body.index("\n}\n")]
    first = float(re.search(r"double q = (\S+);", body).group(1))
    rest = [float(v) for v in re.findall(r"q = fma\(q, f2, (\S+)\);", body)]
    return [first] + rest


def test_log_polynomial_on_the_reduced_range():
    coef = log_coefficients()
    assert len(coef) == 8
    fmax = (np.sqrt(2.0) - 1) / (np.sqrt(2.0) + 1)
    f = np.linspace(0.0, fmax, 100001).astype(np.longdouble)
    x = f * f
    # q(x) = sum_k x^k / (2k + 3), summed in extended precision (x <= 0.0295: 30 terms)
    exact = sum(x ** k / np.longdouble(2 * k + 3) for k in range(30))
    q = np.full_like(x, coef[0])
    for c in coef[1:]:
        q = q * x + np.longdouble(c)
    # q's relative error enters the logarithm scaled by x q <= 0.01: 5e-16 here is 0.05 ulp
    assert float(np.max(np.abs(q - exact) / exact)) <= 5e-16
