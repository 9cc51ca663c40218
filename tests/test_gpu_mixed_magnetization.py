"""Magnetization kind (OMPMagnetizationImp, jif3D/Magnetics/OMPMagnetizationImp.cpp:25-117)
in mixed precision: B = 1e9 mu0/(4 pi G) T M with the tensor far field evaluated once per
site-cell pair for the three components Mx, My, Mz, and the gradient as three tensor
adjoints.

What pins it:
* the FP64 magnetization path (itself pinned to the SimPEG-pinned susceptibility imp through
  M = chi F b_hat 1e-9/mu0, tests/test_oracle_magnetic.py): data within 1e-6 of each
  component's survey maximum for induced-plus-remanent models whose cells do not cancel,
  the gradient within 1e-6 of its largest entry, with and without total-field transforms;
* the mixed susceptibility kind at the same field: with M = chi F b_hat 1e-9/mu0 both mixed
  paths compute the same field from the same tensor terms, to FP32 rounding;
* linearity: the mixed adjoint is the transpose of the mixed forward, <J m, r> = <m, J^T r>/2;
* batches equal single calls bit for bit."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.imp import (HipMagnetizationImp, MinMemGravMagCalculator, TotalField,
                           TotalFieldAnomaly)
from jif3d_amd.model import (ThreeDMagnetizationModel, ThreeDSusceptibilityModel,
                             TotalFieldMagneticData)

pytestmark = pytest.mark.gpu
INC, DEC, FS = 1.1, 0.2, 50000.0
MU0 = 4e-7 * np.pi


def field_dir(inc=INC, dec=DEC):
    return np.array([np.cos(inc) * np.cos(dec), np.cos(inc) * np.sin(dec), np.sin(inc)])


def geometry(model, cubes, seed, n=(40, 36, 14), nsx=10):
    rng = np.random.default_rng(seed)
    model.SetMeshSize(*n)
    if cubes:
        for f, k in zip((model.SetXCellSizes, model.SetYCellSizes, model.SetZCellSizes), n):
            f(np.full(k, 10.0))
    else:
        model.SetXCellSizes(np.full(n[0], 10.0))
        model.SetYCellSizes(np.full(n[1], 12.0))
        model.SetZCellSizes(5.0 * 1.1 ** np.arange(n[2]))
    d = TotalFieldMagneticData()
    xl, yl = sum(model.sizes[0]), sum(model.sizes[1])
    for i in range(nsx):
        for j in range(nsx):
            d.AddMeasurementPoint((i + 0.4) * xl / nsx, (j + 0.7) * yl / nsx,
                                  -rng.uniform(1, 20))
    return d, rng


def induced_magnetization(rng, n):
    """Induced magnetization of positive susceptibility plus a weaker remanent part."""
    chi = rng.random(n) * 0.05
    M = np.concatenate([chi * FS * 1e-9 / MU0 * c for c in field_dir()])
    return M + 0.2 * np.abs(M).max() * rng.random(3 * n), chi


def context(cubes, seed=3):
    m = ThreeDMagnetizationModel()
    d, rng = geometry(m, cubes, seed)
    ctx = nat.Context(0)
    ctx.configure(m, d, (INC, DEC, FS))
    return m, d, ctx, rng


@pytest.mark.parametrize("cubes", [True, False])
def test_mixed_matches_fp64(cubes):
    m, d, ctx, rng = context(cubes)
    K = nat.KIND_MAGNETIZATION
    M, _ = induced_magnetization(rng, m.ncells)
    f64 = ctx.forward(K, nat.FP64, M).reshape(-1, 3)
    mix = ctx.forward(K, nat.MIXED, M).reshape(-1, 3)
    scale = np.abs(f64).max(axis=0)
    assert np.all(np.abs(mix - f64) <= 1e-6 * scale)
    r = rng.normal(size=3 * d.nsites)
    g64 = ctx.adjoint(K, nat.FP64, M, r)
    gmx = ctx.adjoint(K, nat.MIXED, M, r)
    assert gmx.shape == (3 * m.ncells,)
    assert np.max(np.abs(gmx - g64)) <= 1e-6 * np.max(np.abs(g64))


@pytest.mark.parametrize("transform", ["total", "tfa"])
def test_mixed_transforms_through_the_imp(transform):
    m, d, _, rng = context(True, seed=5)
    M, _ = induced_magnetization(rng, m.ncells)
    n = m.ncells
    m.SetMagnetization(M[:n], M[n:2 * n], M[2 * n:])
    calcs = []
    for prec in (nat.FP64, nat.MIXED):
        c = MinMemGravMagCalculator(HipMagnetizationImp(prec))
        c.SetDataTransform(TotalField() if transform == "total" else
                           TotalFieldAnomaly(INC, DEC))
        calcs.append(c)
    f64, mix = (c.Calculate(m, d) for c in calcs)
    assert np.max(np.abs(mix - f64)) <= 1e-6 * np.max(np.abs(f64))
    r = rng.normal(size=d.nsites)
    g64, gmx = (c.LQDerivative(m, d, r) for c in calcs)
    assert np.max(np.abs(gmx - g64)) <= 1e-6 * np.max(np.abs(g64))


def test_mixed_equals_mixed_susceptibility_for_induced_magnetization():
    ms = ThreeDSusceptibilityModel()
    d, rng = geometry(ms, True, 7)
    ctx = nat.Context(0)
    ctx.configure(ms, d, (INC, DEC, FS))
    chi = rng.random(ms.ncells) * 0.05
    Bs = ctx.forward(nat.KIND_MAGNETIC, nat.MIXED, chi)
    M = np.concatenate([chi * FS * 1e-9 / MU0 * c for c in field_dir()])
    Bm = ctx.forward(nat.KIND_MAGNETIZATION, nat.MIXED, M)
    scale = np.abs(Bs.reshape(-1, 3)).max(axis=0)
    assert np.all(np.abs(Bm - Bs).reshape(-1, 3) <= 2e-7 * scale)


def test_mixed_adjoint_is_the_transpose():
    m, d, ctx, rng = context(True, seed=9)
    K = nat.KIND_MAGNETIZATION
    M = rng.normal(size=3 * m.ncells)
    r = rng.normal(size=3 * d.nsites)
    y = ctx.forward(K, nat.MIXED, M)
    g = ctx.adjoint(K, nat.MIXED, M, r)
    lhs, rhs = 2 * np.dot(y, r), np.dot(M, g)
    assert abs(lhs - rhs) <= 1e-6 * (2 * np.abs(y).dot(np.abs(r)))


def test_mixed_batch_equals_single():
    m, d, ctx, rng = context(True, seed=11)
    K = nat.KIND_MAGNETIZATION
    Ms = np.stack([induced_magnetization(rng, m.ncells)[0] for _ in range(3)])
    R = rng.normal(size=(3, 3 * d.nsites))
    F = ctx.forward_batch(K, nat.MIXED, Ms)
    G = ctx.adjoint_batch(K, nat.MIXED, Ms, R)
    for b in range(3):
        assert np.array_equal(F[b], ctx.forward(K, nat.MIXED, Ms[b]))
        assert np.array_equal(G[b], ctx.adjoint(K, nat.MIXED, Ms[b], R[b]))
