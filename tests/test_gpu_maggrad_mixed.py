"""The vertical magnetic gradient in mixed precision (slab form: Bz of 1 m slabs at the
mesh planes through the FP32 multipole far field and FP64 near boxes) against the FP64
two-point form of OMPMagneticGradImp (jif3D/Magnetics/OMPMagneticGradImp.cpp:13-106).

Bound as for the other mixed kinds: data within 1e-6 of the survey's largest magnitude,
gradients within 1e-6 of the largest entry.  Sites whose site or 1 m-lower point is on a
mesh line take the two-point FP64 form and agree to FP64 rounding; cells thinner than 1 m
run the FP64 path (equal bit for bit).  The slab form itself is pinned on the CPU in
tests/test_oracle_maggrad_slabs.py."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd import synthetic
from jif3d_amd.imp import HipMagneticGradImp, MinMemGravMagCalculator
from jif3d_amd.model import ThreeDSusceptibilityModel, TotalFieldMagneticData

pytestmark = pytest.mark.gpu
K = nat.KIND_MAGNETIC_GRADIENT
TOL = 1e-6


def rel_inf(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_headline_like_mesh():
    m, d, field = synthetic.config4_magnetic(n=64, nzc=32, nsite=32)
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    mv = m.GetModelParameters()
    ref = ctx.forward(K, nat.FP64, mv)
    mix = ctx.forward(K, nat.MIXED, mv)
    assert 0.0 < rel_inf(mix, ref) <= TOL
    r = np.random.default_rng(3).normal(size=ref.size)
    g64 = ctx.adjoint(K, nat.FP64, mv, r)
    g = ctx.adjoint(K, nat.MIXED, mv, r)
    assert 0.0 < rel_inf(g, g64) <= TOL


def ragged(seed, dz_min=2):
    rng = np.random.default_rng(seed)
    m = ThreeDSusceptibilityModel()
    nx, ny, nz = 30, 26, 14
    m.SetMeshSize(nx, ny, nz)
    m.SetXCellSizes(rng.integers(10, 30, nx) + 0.0)
    m.SetYCellSizes(rng.integers(10, 30, ny) + 0.0)
    dz = rng.integers(dz_min, 20, nz) + 0.0
    m.SetZCellSizes(dz)
    m.SetSusceptibilities(rng.random(m.ncells) * 0.05)
    m.SetBackgroundThicknesses([20.0, 1000.0])
    m.SetBackgroundSusceptibilities([0.01, 0.02])
    X, Y, Z = m.GetXCoordinates(), m.GetYCoordinates(), m.GetZCoordinates()
    d = TotalFieldMagneticData()
    for _ in range(150):
        d.AddMeasurementPoint(rng.uniform(X[0] - 50, X[-1] + 50), rng.uniform(Y[0] - 50, Y[-1] + 50),
                              -rng.uniform(0.5, 60))
    special = [(X[7], Y[5], -10.0),                    # above the mesh on a line
               (X[9], Y[4] + 3.3, Z[0] - 1.0),         # 1 m-lower point on a line
               (X[12] + 4.1, Y[11], Z[3] - 1.0),       # the same inside the mesh
               (X[5], Y[8] + 2.2, Z[4]),               # on a line inside the mesh
               (0.5 * (X[15] + X[16]), 0.5 * (Y[9] + Y[10]), 0.5 * (Z[5] + Z[6])),  # in a cell
               (X[20] + 1.1, Y[20] + 2.3, Z[2] - 1.0)]  # on a slab plane only
    for p in special:
        d.AddMeasurementPoint(*p)
    return m, d, len(special)


@pytest.mark.parametrize("seed", [5, 6])
def test_ragged_mesh_lines_and_inside(seed):
    m, d, nsp = ragged(seed)
    field = (0.7, -0.4, 52000.0)
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    mv = m.GetModelParameters()
    ref = ctx.forward(K, nat.FP64, mv)
    mix = ctx.forward(K, nat.MIXED, mv)
    assert rel_inf(mix, ref) <= TOL
    line = np.arange(d.nsites - nsp, d.nsites - 2)  # the four special sites on lines
    assert np.max(np.abs(mix[line] - ref[line])) <= 1e-12 * np.max(np.abs(ref))
    r = np.random.default_rng(seed).normal(size=ref.size)
    g64 = ctx.adjoint(K, nat.FP64, mv, r)
    g = ctx.adjoint(K, nat.MIXED, mv, r)
    assert rel_inf(g, g64) <= TOL
    assert np.all(g[m.ncells:] == 0.0)
    # batches: each model of the batch as alone
    M = np.stack([mv, 1.5 * mv, mv[::-1].copy()])
    F = ctx.forward_batch(K, nat.MIXED, M)
    R = np.stack([r, -r, 2 * r])
    Gb = ctx.adjoint_batch(K, nat.MIXED, M, R)
    for b in range(3):
        assert np.array_equal(F[b], ctx.forward(K, nat.MIXED, M[b]))
        assert np.array_equal(Gb[b], ctx.adjoint(K, nat.MIXED, M[b], R[b]))


def test_one_metre_cells():
    """Cells exactly 1 m thick leave no gap layer: their two slabs are adjacent layers of the
    slab mesh, both in the far field's layer list."""
    m, d, _ = ragged(11)
    dz = m.GetZCellSizes().copy()
    dz[2] = dz[3] = dz[9] = 1.0
    m.SetZCellSizes(dz)
    ctx = nat.Context(0)
    ctx.configure(m, d, (0.2, 1.1, 49000.0))
    mv = m.GetModelParameters()
    ref = ctx.forward(K, nat.FP64, mv)
    assert rel_inf(ctx.forward(K, nat.MIXED, mv), ref) <= TOL
    r = np.random.default_rng(12).normal(size=ref.size)
    assert rel_inf(ctx.adjoint(K, nat.MIXED, mv, r), ctx.adjoint(K, nat.FP64, mv, r)) <= TOL


def test_thin_cells_run_fp64():
    m, d, _ = ragged(7, dz_min=2)
    dz = m.GetZCellSizes().copy()
    dz[3] = 0.5
    m.SetZCellSizes(dz)
    ctx = nat.Context(0)
    ctx.configure(m, d, (0.7, -0.4, 52000.0))
    mv = m.GetModelParameters()
    assert np.array_equal(ctx.forward(K, nat.MIXED, mv), ctx.forward(K, nat.FP64, mv))
    r = np.random.default_rng(8).normal(size=d.nsites)
    assert np.array_equal(ctx.adjoint(K, nat.MIXED, mv, r), ctx.adjoint(K, nat.FP64, mv, r))


def test_imp_mixed_precision_and_geometry_change():
    m, d, field = synthetic.config4_magnetic(n=48, nzc=24, nsite=24)
    ref = MinMemGravMagCalculator(HipMagneticGradImp(*field))
    mixed = MinMemGravMagCalculator(HipMagneticGradImp(*field, precision=nat.MIXED))
    assert rel_inf(mixed.Calculate(m, d), ref.Calculate(m, d)) <= TOL
    r = np.random.default_rng(9).normal(size=d.nsites)
    assert rel_inf(mixed.LQDerivative(m, d, r), ref.LQDerivative(m, d, r)) <= TOL
    # a new survey and a new field on the same calculator
    d2 = TotalFieldMagneticData()
    for i in range(20):
        d2.AddMeasurementPoint(15.0 + 37.0 * i, 400.0 + 11.0 * i, -20.0 - i)
    f2 = (0.3, 0.9, 48000.0)
    ref2 = MinMemGravMagCalculator(HipMagneticGradImp(*f2))
    mixed2 = MinMemGravMagCalculator(HipMagneticGradImp(*f2, precision=nat.MIXED))
    assert rel_inf(mixed2.Calculate(m, d2), ref2.Calculate(m, d2)) <= TOL
    assert rel_inf(mixed.Calculate(m, d2), ref.Calculate(m, d2)) <= TOL


def test_edge_cases():
    """No cells (background only: the vertical gradient has no background terms), one cell
    and one site (inside every near box: FP64 throughout), no sites."""
    field = (0.7, -0.4, 52000.0)
    m = ThreeDSusceptibilityModel()
    m.SetMeshSize(0, 0, 0)
    m.SetBackgroundThicknesses([200.0, 3500.0])
    m.SetBackgroundSusceptibilities([0.01, 0.02])
    d = TotalFieldMagneticData()
    d.SetMeasurementPoints([3e4, 5e4], [2e4, 6e4], [0.0, -10.0])
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    mv = m.GetModelParameters()
    assert np.array_equal(ctx.forward(K, nat.MIXED, mv), ctx.forward(K, nat.FP64, mv))
    r = np.array([1.0, -2.0])
    assert np.array_equal(ctx.adjoint(K, nat.MIXED, mv, r), ctx.adjoint(K, nat.FP64, mv, r))
    m = ThreeDSusceptibilityModel()
    m.SetMeshSize(1, 1, 1)
    for f in (m.SetXCellSizes, m.SetYCellSizes, m.SetZCellSizes):
        f([20.0])
    m.SetSusceptibilities([0.05])
    d = TotalFieldMagneticData()
    d.AddMeasurementPoint(5.0, 7.0, -3.0)
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    mv = m.GetModelParameters()
    ref = ctx.forward(K, nat.FP64, mv)
    assert rel_inf(ctx.forward(K, nat.MIXED, mv), ref) <= 1e-12
    g64 = ctx.adjoint(K, nat.FP64, mv, np.array([1.5]))
    assert rel_inf(ctx.adjoint(K, nat.MIXED, mv, np.array([1.5])), g64) <= 1e-12
    imp = HipMagneticGradImp(*field, precision=nat.MIXED)
    assert imp.Calculate(m, TotalFieldMagneticData()).size == 0
