"""FullSensitivityGravMagCalculator on the GPU: J kept in HBM (FP32), cached forward
and gradient through the C ABI, following the reference's cached-calculator tests
(jif3D/Gravity/test_ScalarCalculation.cpp:248-305, test_GravityObjective.cpp:31-110).

Tolerances differ from the reference on purpose: its J is FP64 (float epsilon
agreement between direct and cached results), this J is FP32, so cached results
agree with the direct FP64 path to 2e-6 of each component's largest value and
gradients to 2e-6 of ||g||_inf."""
import time

import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.imp import (ScalarHipGravityImp, TensorHipGravityImp, ScalarTensorHipGravityImp,
                           HipMagneticSusceptibilityImp, FullSensitivityGravMagCalculator,
                           MinMemGravMagCalculator, FTGInvariant, TotalFieldAnomaly)
from jif3d_amd.model import ThreeDGravityModel, TensorGravityData

pytestmark = pytest.mark.gpu


def per_component_rel(a, b, k):
    a, b = a.reshape(-1, k), b.reshape(-1, k)
    floor = 1e-6 * np.max(np.abs(b))
    return max(np.max(np.abs(a[:, c] - b[:, c])) / max(np.max(np.abs(b[:, c])), floor)
               for c in range(k))


def rel_inf(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def random_model(seed, nx=9, ny=8, nz=6, nmeas=25, bg=True, scale=1.0):
    rng = np.random.default_rng(seed)
    m = ThreeDGravityModel()
    m.SetMeshSize(nx, ny, nz)
    m.SetXCellSizes(rng.integers(10, 30, nx) + 0.0)
    m.SetYCellSizes(rng.integers(10, 30, ny) + 0.0)
    m.SetZCellSizes(rng.integers(5, 20, nz) + 0.0)
    m.SetDensities(scale * (1.0 + 10.0 * rng.random(nx * ny * nz)))
    if bg:
        m.SetBackgroundThicknesses([15.0, 200.0, 1e4])
        m.SetBackgroundDensities([2.0 * scale, 3.0 * scale, 1.5 * scale])
    d = TensorGravityData()
    xl, yl = sum(m.sizes[0]), sum(m.sizes[1])
    for _ in range(nmeas):
        d.AddMeasurementPoint(rng.uniform(0, xl), rng.uniform(0, yl), -rng.uniform(1, 60))
    return m, d


IMPS = {
    "scalar": (lambda: ScalarHipGravityImp(), 1),
    "tensor": (lambda: TensorHipGravityImp(), 9),
    "scalar_tensor": (lambda: ScalarTensorHipGravityImp(), 10),
    "magnetic": (lambda: HipMagneticSusceptibilityImp(1.1, 0.2, 50000.0), 3),
}


@pytest.mark.parametrize("name", list(IMPS))
def test_cached_forward_matches_direct(name):
    make, k = IMPS[name]
    m, d = random_model(1, scale=1e-3 if name == "magnetic" else 1.0)
    calc = FullSensitivityGravMagCalculator(make())
    first = calc.Calculate(m, d)          # direct (FP64) and J built
    second = calc.Calculate(m, d)         # J m through the stored J
    assert first.shape == second.shape == (d.nsites * k,)
    assert per_component_rel(second, first, k) <= 2e-6
    # a new model on the same geometry stays cached and still agrees
    rng = np.random.default_rng(2)
    dens = m.GetModelParameters().copy()
    dens *= 1.0 + 0.2 * rng.random(dens.size)
    m.SetDensities(dens[:m.ncells])
    m.SetBackgroundDensities(dens[m.ncells:])
    ref = MinMemGravMagCalculator(make()).Calculate(m, d)
    assert per_component_rel(calc.Calculate(m, d), ref, k) <= 2e-6


def test_cache_is_faster_and_invalidated_by_geometry():
    m, d = random_model(3, nx=32, ny=32, nz=16, nmeas=400)
    calc = FullSensitivityGravMagCalculator(ScalarTensorHipGravityImp())
    t0 = time.perf_counter()
    calc.Calculate(m, d)
    t1 = time.perf_counter()
    calc.Calculate(m, d)
    t2 = time.perf_counter()
    assert t1 - t0 > t2 - t1
    # moving one site changes the geometry: the next call is a direct calculation
    x, y, z = d.positions
    d2 = TensorGravityData()
    for i in range(d.nsites):
        d2.AddMeasurementPoint(x[i] + (5.0 if i == 0 else 0.0), y[i], z[i])
    ref = MinMemGravMagCalculator(ScalarTensorHipGravityImp()).Calculate(m, d2)
    got = calc.Calculate(m, d2)
    assert per_component_rel(got, ref, 10) <= 1e-12
    assert per_component_rel(calc.Calculate(m, d2), ref, 10) <= 2e-6


@pytest.mark.parametrize("name", list(IMPS))
def test_cached_lqderivative(name, orc):
    """LQDerivative before the cache (matrix free), after it (2 [J B]^T r through the
    stored J) and 2 J^T r from the oracle's FP64 rows agree (lqderivative_test)."""
    make, k = IMPS[name]
    m, d = random_model(4, scale=1e-3 if name == "magnetic" else 1.0)
    calc = FullSensitivityGravMagCalculator(make())
    misfit = np.random.default_rng(5).random(d.nsites * k)
    deriv = calc.LQDerivative(m, d, misfit)
    calc.Calculate(m, d)
    deriv2 = calc.LQDerivative(m, d, misfit)
    okind = {"scalar": orc.SCALAR, "tensor": orc.TENSOR, "magnetic": orc.MAGNETIC}
    if name == "scalar_tensor":
        r = misfit.reshape(-1, 10)
        cmp = (orc.lqderivative(orc.SCALAR, m, m.GetModelParameters(), d.positions,
                                r[:, 0].copy())
               + orc.lqderivative(orc.TENSOR, m, m.GetModelParameters(), d.positions,
                                  r[:, 1:].reshape(-1).copy()))
    else:
        field = (1.1, 0.2, 50000.0) if name == "magnetic" else None
        cmp = orc.lqderivative(okind[name], m, m.GetModelParameters(), d.positions, misfit,
                               field=field)
    assert deriv.shape == deriv2.shape == cmp.shape == (m.ncells + 3,)
    assert rel_inf(deriv, cmp) <= 1e-10
    assert rel_inf(deriv2, cmp) <= 2e-6


@pytest.mark.parametrize("tf", ["ftg_invariant", "tfa"])
def test_cached_transforms(tf):
    """Transformed data and the gradient through the transform Jacobian at the cached
    raw data (CachedLQDerivative, FullSensitivityGravMagCalculator.h:166-211)."""
    if tf == "ftg_invariant":
        make, transform, scale = (lambda: TensorHipGravityImp()), FTGInvariant(), 1.0
    else:
        make = lambda: HipMagneticSusceptibilityImp(1.1, 0.2, 50000.0)
        transform, scale = TotalFieldAnomaly(1.0, 0.3), 1e-3
    m, d = random_model(6, scale=scale)
    calc = FullSensitivityGravMagCalculator(make())
    calc.SetDataTransform(transform)
    direct = MinMemGravMagCalculator(make())
    direct.SetDataTransform(transform)
    first = calc.Calculate(m, d)
    second = calc.Calculate(m, d)
    ref = direct.Calculate(m, d)
    assert rel_inf(first, ref) <= 1e-12
    assert rel_inf(second, ref) <= 5e-6
    misfit = np.random.default_rng(7).normal(size=d.nsites)
    g = calc.LQDerivative(m, d, misfit)
    gref = direct.LQDerivative(m, d, misfit)
    assert rel_inf(g, gref) <= 5e-6


@pytest.mark.parametrize("prec", [nat.GEMM_BF16, nat.GEMM_BF16_REFINED])
def test_cached_low_precision_modes(prec):
    m, d = random_model(8)
    calc = FullSensitivityGravMagCalculator(ScalarTensorHipGravityImp(), gemm_precision=prec)
    first = calc.Calculate(m, d)
    second = calc.Calculate(m, d)
    tol = 2e-2 if prec == nat.GEMM_BF16 else 1e-4
    assert per_component_rel(second, first, 10) <= tol
