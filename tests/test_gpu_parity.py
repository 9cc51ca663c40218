"""GPU parity of the FP64 path against the CPU oracle, through the C ABI.

Tolerances: the FP64 kernels reorder the reference's sums (corner-first
accumulation) and use the GPU's log/atan2, so results agree to rounding, not bit
for bit: per-site data within 1e-12 relative (gz) or 1e-11 of the component's
largest magnitude (tensor, magnetic); gradients within 1e-10 of ||g||_inf."""
import os

import numpy as np
import pytest

from golden_paths import GOLDEN
from jif3d_amd import _native as nat
from jif3d_amd.imp import (ScalarHipGravityImp, TensorHipGravityImp, ScalarTensorHipGravityImp,
                           HipMagneticSusceptibilityImp, MinMemGravMagCalculator, FTGInvariant,
                           TotalFieldAnomaly, TotalField)
from jif3d_amd.model import (ThreeDGravityModel, ThreeDSusceptibilityModel, ScalarGravityData,
                             TensorGravityData)
from jif3d_amd import synthetic

pytestmark = pytest.mark.gpu
G = 6.67428e-11


def rel_inf(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def per_component_rel(a, b, k):
    """Max error of each component over its largest magnitude; a component that is
    zero by symmetry (rounding noise only) is measured against 1e-6 of the largest
    component instead of its own noise."""
    a = a.reshape(-1, k)
    b = b.reshape(-1, k)
    floor = 1e-6 * max(np.max(np.abs(b)), 1e-300)
    return max(np.max(np.abs(a[:, c] - b[:, c])) / max(np.max(np.abs(b[:, c])), floor)
               for c in range(k))


def random_model(rng, maxcells, nmeas, bg=True):
    m = ThreeDGravityModel()
    dims = [int(rng.integers(0, maxcells)) + 1 for _ in range(3)]
    m.SetMeshSize(*dims)
    sizes = [rng.integers(0, 25, d) + 10.0 for d in dims]
    m.SetXCellSizes(sizes[0])
    m.SetYCellSizes(sizes[1])
    m.SetZCellSizes(sizes[2])
    m.SetDensities(rng.random(int(np.prod(dims))) * 10 + 1.0)
    if bg:
        m.SetBackgroundThicknesses([15.0, 200.0, 1e4])
        m.SetBackgroundDensities([2.0, 3.0, 1.5])
    d = TensorGravityData()
    xl, yl = int(np.floor(sizes[0].sum())), int(np.floor(sizes[1].sum()))
    for _ in range(nmeas):
        d.AddMeasurementPoint(rng.integers(0, xl) + 1.5, rng.integers(0, yl) + 0.5,
                              -(rng.integers(0, 100) + 1.0))
    return m, d


def test_config0_scalar_forward(orc):
    """Config 0: 20x20x10 mesh, 100 sites; gz within 1e-12 relative of the oracle."""
    m, d = synthetic.config0()
    imp = ScalarHipGravityImp()
    gpu = MinMemGravMagCalculator(imp).Calculate(m, d)
    ref = orc.calculate(orc.SCALAR, m, m.GetModelParameters(), d.positions)
    assert np.max(np.abs(gpu - ref) / np.abs(ref)) <= 1e-12
    truth = orc.calculate_ld(orc.SCALAR, m, m.GetModelParameters(), d.positions)
    assert np.max(np.abs(gpu - truth) / np.abs(truth)) <= 1e-12


def test_single_prism_known_answers():
    """Li & Chouteau values through the GPU (test_ScalarCalculation.cpp:32-64)."""
    m = ThreeDGravityModel()
    m.SetMeshSize(1, 1, 1)
    for ax in (m.SetXCellSizes, m.SetYCellSizes, m.SetZCellSizes):
        ax([20.0])
    m.SetOrigin(-10.0, -10.0, -10.0)
    m.SetDensities([1.0])
    d = ScalarGravityData()
    for p in [(0, 0, 0), (0, 0, -10), (0, 0, 10), (1e5, 10, 10), (100, 100, 10)]:
        d.AddMeasurementPoint(*p)
    g = ScalarHipGravityImp().Calculate(m, d)
    assert abs(g[0]) < np.finfo(float).eps
    assert abs(g[1] - 3.46426e-9) <= 1e-3 * 3.46426e-9
    assert abs(g[1] + g[2]) <= 1.2e-9 * abs(g[1])
    assert g[3] < np.finfo(float).eps
    assert abs(g[4] + 1.873178e-12) <= 1e-3 * 1.873178e-12


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_random_models_all_kinds(orc, seed):
    rng = np.random.default_rng(100 + seed)
    m, d = random_model(rng, 12, 9)
    mv = m.GetModelParameters()
    for imp, okind, k in [(ScalarHipGravityImp(), orc.SCALAR, 1),
                          (TensorHipGravityImp(), orc.TENSOR, 9)]:
        gpu = imp.Calculate(m, d)
        ref = orc.calculate(okind, m, mv, d.positions)
        tol = 1e-12 if k == 1 else 1e-11
        assert per_component_rel(gpu, ref, k) <= tol, (seed, k)
    both = ScalarTensorHipGravityImp().Calculate(m, d).reshape(-1, 10)
    ref_g = orc.calculate(orc.SCALAR, m, mv, d.positions)
    ref_t = orc.calculate(orc.TENSOR, m, mv, d.positions)
    assert rel_inf(both[:, 0], ref_g) <= 1e-12
    assert per_component_rel(both[:, 1:].ravel(), ref_t, 9) <= 1e-11


def test_tensor_inside_body_poisson(orc):
    """Sites inside the mesh: trace = -4 pi G rho (test_TensorCalculation.cpp:187-227)."""
    m = ThreeDGravityModel()
    m.SetMeshSize(10, 10, 10)
    for ax in (m.SetXCellSizes, m.SetYCellSizes, m.SetZCellSizes):
        ax(np.full(10, 10.0))
    rho = 2.1
    m.SetDensities(np.full(1000, rho))
    m.SetBackgroundThicknesses([100.0])
    m.SetBackgroundDensities([rho])
    d = TensorGravityData()
    d.AddMeasurementPoint(5, 5, -1)
    d.AddMeasurementPoint(5, 5, 50)
    t = TensorHipGravityImp().Calculate(m, d)
    tr1, tr2 = t[0] + t[4] + t[8], t[9] + t[13] + t[17]
    assert abs((tr1 - tr2) + 4 * np.pi * G * rho) <= 1.2e-9 * 4 * np.pi * G * rho
    ref = orc.calculate(orc.TENSOR, m, m.GetModelParameters(), d.positions)
    assert rel_inf(t, ref) <= 1e-11  # site 1 lies above a uniform slab: its tensor is noise


def test_magnetic_simpeg_golden(orc):
    m = ThreeDSusceptibilityModel()
    m.SetMeshSize(10, 10, 10)
    m.SetXCellSizes(np.full(10, 2.0))
    m.SetYCellSizes(np.full(10, 4.0))
    m.SetZCellSizes(np.full(10, 0.5))
    m.SetOrigin(-10.0, -20.0, 0.0)
    m.SetSusceptibilities(np.full(1000, 0.2))
    d = TensorGravityData()
    d.SetMeasurementPoints(-20.0 + np.arange(200) / 5.0, np.full(200, 10.0), np.full(200, -1.0))
    inc, dec, fs = np.radians(15.0), np.radians(10.0), 50000.0
    imp = HipMagneticSusceptibilityImp(inc, dec, fs)
    b = MinMemGravMagCalculator(imp).Calculate(m, d).reshape(-1, 3)
    for comp, name in enumerate(("magx.simpeg", "magy.simpeg", "magz.simpeg")):
        np.testing.assert_allclose(b[:, comp], np.loadtxt(os.path.join(GOLDEN, name)), rtol=2e-3)
    ref = orc.calculate(orc.MAGNETIC, m, m.GetModelParameters(), d.positions, field=(inc, dec, fs))
    assert per_component_rel(b.ravel(), ref, 3) <= 1e-11
    calc = MinMemGravMagCalculator(imp)
    calc.SetDataTransform(TotalFieldAnomaly(inc, dec))
    tfa = calc.Calculate(m, d)
    np.testing.assert_allclose(tfa, np.loadtxt(os.path.join(GOLDEN, "magt.simpeg")), rtol=2e-3)


@pytest.mark.parametrize("which", ["scalar", "tensor", "magnetic", "ftginv", "tfa", "tf"])
def test_adjoint_matches_oracle(orc, which):
    rng = np.random.default_rng(7)
    m, d = random_model(rng, 9, 8)
    field = (0.4, 0.3, 50000.0)
    mv = m.GetModelParameters()
    spec = {
        "scalar": (ScalarHipGravityImp(), orc.SCALAR, None, orc.TF_NONE, None),
        "tensor": (TensorHipGravityImp(), orc.TENSOR, None, orc.TF_NONE, None),
        "magnetic": (HipMagneticSusceptibilityImp(*field), orc.MAGNETIC, field, orc.TF_NONE, None),
        "ftginv": (TensorHipGravityImp(), orc.TENSOR, None, orc.TF_FTG_INVARIANT, FTGInvariant()),
        "tfa": (HipMagneticSusceptibilityImp(*field), orc.MAGNETIC, field,
                orc.TF_TOTAL_FIELD_ANOMALY, TotalFieldAnomaly(0.2, 0.1)),
        "tf": (HipMagneticSusceptibilityImp(*field), orc.MAGNETIC, field, orc.TF_TOTAL_FIELD,
               TotalField()),
    }[which]
    imp, okind, fld, otf, tf = spec
    calc = MinMemGravMagCalculator(imp)
    angles = (0.2, 0.1) if otf == orc.TF_TOTAL_FIELD_ANOMALY else (0.0, 0.0)
    if tf is not None:
        calc.SetDataTransform(tf)
    fwd = calc.Calculate(m, d)
    ref_fwd = orc.calculate(okind, m, mv, d.positions, field=fld, transform=otf, tf_angles=angles)
    assert rel_inf(fwd, ref_fwd) <= 1e-11
    misfit = rng.normal(size=fwd.size)
    g = calc.LQDerivative(m, d, misfit)
    ref = orc.lqderivative(okind, m, mv, d.positions, misfit, field=fld, transform=otf,
                           tf_angles=angles)
    assert g.shape == ref.shape
    assert rel_inf(g, ref) <= 1e-10, which


def test_sensitivity_rows_match_oracle(orc):
    rng = np.random.default_rng(9)
    m, d = random_model(rng, 7, 5)
    mv = m.GetModelParameters()
    for imp, okind in [(ScalarHipGravityImp(), orc.SCALAR), (TensorHipGravityImp(), orc.TENSOR)]:
        rows = imp.SensitivityRows(m, d)
        _, J = orc.calculate(okind, m, mv, d.positions, want_sens=True)
        assert rows.shape == J.shape
        scale = np.max(np.abs(J), axis=1, keepdims=True)
        assert np.max(np.abs(rows - J) / scale) <= 1e-12


def test_edge_cases(orc):
    # background only (no cells), sites at the surface
    m = ThreeDGravityModel()
    m.SetMeshSize(0, 0, 0)
    m.SetBackgroundThicknesses([200.0, 300.0, 3500.0, 1000.0])
    m.SetBackgroundDensities([1.0, 1.0, 5.0, 5.0])
    d = ScalarGravityData()
    d.SetMeasurementPoints([3e4, 5e4], [2e4, 6e4], [0.0, 0.0])
    g = ScalarHipGravityImp().Calculate(m, d)
    analytic = 2 * np.pi * G * (500.0 + 5.0 * 4500.0)
    np.testing.assert_allclose(g, analytic, rtol=1e-4)
    # one cell, one site, ragged sizes
    m = ThreeDGravityModel()
    m.SetMeshSize(1, 3, 2)
    m.SetXCellSizes([7.0])
    m.SetYCellSizes([1.0, 13.0, 2.5])
    m.SetZCellSizes([3.0, 40.0])
    m.SetDensities(np.arange(6) + 1.0)
    d = ScalarGravityData()
    d.AddMeasurementPoint(2.0, 3.0, -0.5)
    gpu = ScalarHipGravityImp().Calculate(m, d)
    ref = orc.calculate(orc.SCALAR, m, m.GetModelParameters(), d.positions)
    assert rel_inf(gpu, ref) <= 1e-12
    # no sites
    d0 = ScalarGravityData()
    assert ScalarHipGravityImp().Calculate(m, d0).size == 0
    g0 = ScalarHipGravityImp().LQDerivative(m, d0, np.zeros(0))
    assert np.all(g0 == 0) and g0.size == 6


def test_repeat_runs_are_bitwise_identical():
    """The reference's CUDA test reruns 10 times and demands identical output."""
    rng = np.random.default_rng(5)
    m, d = random_model(rng, 12, 20)
    imp = TensorHipGravityImp()
    first = imp.Calculate(m, d)
    for _ in range(5):
        assert np.array_equal(imp.Calculate(m, d), first)


def test_errors_are_reported():
    m, d = random_model(np.random.default_rng(1), 4, 3)
    imp = ScalarHipGravityImp()
    calc = MinMemGravMagCalculator(imp)
    calc.SetDataTransform(FTGInvariant())  # invariant needs tensor data
    with pytest.raises(nat.JgmError) as e:
        calc.Calculate(m, d)
    assert "tensor" in str(e.value)
