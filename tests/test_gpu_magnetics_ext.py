"""Magnetic vertical-gradient and magnetization imps on the GPU (FP64) against the
oracle, through the C ABI and the Python imp mirrors.

The oracle's restatements of OMPMagneticGradImp and OMPMagnetizationImp are pinned to
the SimPEG-pinned susceptibility imp (tests/test_oracle_magnetic.py).  Tolerances as
for the other FP64 kinds: data within 1e-11 of the largest magnitude, gradients within
1e-10 of ||g||_inf (corner-first reordering and the GPU's log/atan2)."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.imp import (HipMagneticGradImp, HipMagnetizationImp, MinMemGravMagCalculator,
                           TotalFieldAnomaly, TotalField)
from jif3d_amd.model import (ThreeDSusceptibilityModel, ThreeDMagnetizationModel,
                             TotalFieldMagneticData)

pytestmark = pytest.mark.gpu
FIELD = (0.7, -0.4, 52000.0)


def rel_inf(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def mesh(model, rng, nx=8, ny=7, nz=5):
    model.SetMeshSize(nx, ny, nz)
    model.SetXCellSizes(rng.integers(10, 30, nx) + 0.0)
    model.SetYCellSizes(rng.integers(10, 30, ny) + 0.0)
    model.SetZCellSizes(rng.integers(5, 20, nz) + 0.0)


def sites(model, rng, n=30):
    d = TotalFieldMagneticData()
    xl, yl = sum(model.sizes[0]), sum(model.sizes[1])
    for _ in range(n):
        d.AddMeasurementPoint(rng.uniform(-20, xl + 20), rng.uniform(-20, yl + 20),
                              -rng.uniform(0.5, 40))
    return d


def susceptibility_model(seed, bg=True):
    rng = np.random.default_rng(seed)
    m = ThreeDSusceptibilityModel()
    mesh(m, rng)
    m.SetSusceptibilities(rng.random(m.ncells) * 0.05)
    if bg:
        m.SetBackgroundThicknesses([20.0, 1000.0])
        m.SetBackgroundSusceptibilities([0.01, 0.02])
    return m, sites(m, rng)


def magnetization_model(seed):
    rng = np.random.default_rng(seed)
    m = ThreeDMagnetizationModel()
    mesh(m, rng)
    n = m.ncells
    m.SetMagnetization(rng.normal(size=n), rng.normal(size=n), rng.normal(size=n) + 0.5)
    return m, sites(m, rng)


def test_magnetic_gradient_forward_and_gradient(orc):
    m, d = susceptibility_model(1)
    imp = HipMagneticGradImp(*FIELD)
    calc = MinMemGravMagCalculator(imp)
    mv = m.GetModelParameters()
    got = calc.Calculate(m, d)
    ref = orc.calculate(orc.MAGGRAD, m, mv, d.positions, field=FIELD)
    assert got.shape == (d.nsites,)
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))
    r = np.random.default_rng(2).normal(size=d.nsites)
    g = calc.LQDerivative(m, d, r)
    gref = orc.lqderivative(orc.MAGGRAD, m, mv, d.positions, r, field=FIELD)
    assert g.shape == (m.ncells + 2,)
    assert np.all(g[m.ncells:] == 0.0)  # OMPMagneticGradImp has no background terms
    assert np.max(np.abs(g - gref)) <= 1e-10 * np.max(np.abs(gref))


def test_magnetic_gradient_equals_bz_difference():
    """Through the GPU's own susceptibility imp: 4 pi 1e-7 (Bz(z) - Bz(z + 1))."""
    m, d = susceptibility_model(3, bg=False)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    mv = m.GetModelParameters()
    g = ctx.forward(nat.KIND_MAGNETIC_GRADIENT, nat.FP64, mv)
    b0 = ctx.forward(nat.KIND_MAGNETIC, nat.FP64, mv).reshape(-1, 3)
    x, y, z = d.positions
    ctx.set_sites(x, y, z + 1.0)
    b1 = ctx.forward(nat.KIND_MAGNETIC, nat.FP64, mv).reshape(-1, 3)
    ref = 4 * np.pi * 1e-7 * (b0[:, 2] - b1[:, 2])
    assert np.max(np.abs(g - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("transform", [None, "tfa", "total"])
def test_magnetization_forward_and_gradient(orc, transform):
    m, d = magnetization_model(4)
    imp = HipMagnetizationImp()
    calc = MinMemGravMagCalculator(imp)
    tf, angles = orc.TF_NONE, (0.0, 0.0)
    if transform == "tfa":
        calc.SetDataTransform(TotalFieldAnomaly(0.9, 0.3))
        tf, angles = orc.TF_TOTAL_FIELD_ANOMALY, (0.9, 0.3)
    elif transform == "total":
        calc.SetDataTransform(TotalField())
        tf = orc.TF_TOTAL_FIELD
    mv = m.GetModelParameters()
    got = calc.Calculate(m, d)
    ref = orc.calculate(orc.MAGNETIZATION, m, mv, d.positions, transform=tf, tf_angles=angles)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))
    r = np.random.default_rng(5).normal(size=got.size)
    g = calc.LQDerivative(m, d, r)
    gref = orc.lqderivative(orc.MAGNETIZATION, m, mv, d.positions, r, transform=tf,
                            tf_angles=angles)
    assert g.shape == (3 * m.ncells,)
    assert np.max(np.abs(g - gref)) <= 1e-10 * np.max(np.abs(gref))


def test_magnetization_dot_product():
    """<J m, r> = <m, J^T r>: forward and adjoint are transposes of each other."""
    m, d = magnetization_model(6)
    ctx = nat.Context(0)
    ctx.configure(m, d)
    rng = np.random.default_rng(7)
    mv = rng.normal(size=3 * m.ncells)
    r = rng.normal(size=3 * d.nsites)
    y = ctx.forward(nat.KIND_MAGNETIZATION, nat.FP64, mv)
    g = ctx.adjoint(nat.KIND_MAGNETIZATION, nat.FP64, mv, r)
    assert abs(2 * np.dot(y, r) - np.dot(mv, g)) <= 1e-12 * np.abs(mv).dot(np.abs(g))


def test_model_sizes_of_the_magnetic_extensions():
    m, d = susceptibility_model(8)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    assert ctx.model_size(nat.KIND_MAGNETIZATION) == 3 * m.ncells
    assert ctx.model_size(nat.KIND_MAGNETIC_GRADIENT) == m.ncells + 2
