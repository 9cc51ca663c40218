"""Stored J of the magnetic vertical-gradient and magnetization kinds on the GPU, through
the C ABI and the FullSensitivity / Disk calculator mirrors.

jif3D keeps the sensitivities of any imp in FullSensitivityGravMagCalculator
(jif3D/GravMag/FullSensitivityGravMagCalculator.h:99-221), including those of
OMPMagneticGradImp (one row per site, N cell columns plus zero background columns,
jif3D/Magnetics/OMPMagneticGradImp.cpp:13-106) and OMPMagnetizationImp (three rows per
site, 3N columns [Mx; My; Mz], jif3D/Magnetics/OMPMagnetizationImp.cpp:25-117).

What pins it:
* J itself: every entry against the oracle's sensitivity matrix in the reference layout,
  to FP32 rounding of the FP64 value (the vertical-gradient entry is the FP64 difference
  of the two heights, rounded once), read back exactly as J * I, in HBM and host-resident.
* Products: J m and 2 J^T r against the FP64 matrix-free path within the a-priori bound
  of FP32 products, (K + 2) 2^-24 (|J| |m|) entrywise, K the length of one sum.
* Calculators: the cached result and gradient of FullSensitivityGravMagCalculator agree
  with the direct imp, batches equal single calls bit for bit."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.imp import (HipMagneticGradImp, HipMagnetizationImp, MinMemGravMagCalculator,
                           FullSensitivityGravMagCalculator, DiskGravMagCalculator, TotalField)
from jif3d_amd.model import (ThreeDSusceptibilityModel, ThreeDMagnetizationModel,
                             TotalFieldMagneticData)

pytestmark = pytest.mark.gpu
FIELD = (0.7, -0.4, 52000.0)
KINDS = [nat.KIND_MAGNETIC_GRADIENT, nat.KIND_MAGNETIZATION]


def make_model(kind, seed, nx=9, ny=8, nz=5, nsites=27):
    rng = np.random.default_rng(seed)
    if kind == nat.KIND_MAGNETIZATION:
        m = ThreeDMagnetizationModel()
    else:
        m = ThreeDSusceptibilityModel()
    m.SetMeshSize(nx, ny, nz)
    m.SetXCellSizes(rng.integers(10, 30, nx) + 0.0)
    m.SetYCellSizes(rng.integers(10, 30, ny) + 0.0)
    m.SetZCellSizes(rng.integers(5, 20, nz) + 0.0)
    n = m.ncells
    if kind == nat.KIND_MAGNETIZATION:
        m.SetMagnetization(rng.normal(size=n), rng.normal(size=n), rng.normal(size=n) + 0.5)
    else:
        m.SetSusceptibilities(rng.random(n) * 0.05)
        m.SetBackgroundThicknesses([20.0, 1000.0])
        m.SetBackgroundSusceptibilities([0.01, 0.02])
    d = TotalFieldMagneticData()
    xl, yl = sum(m.sizes[0]), sum(m.sizes[1])
    for _ in range(nsites):
        d.AddMeasurementPoint(rng.uniform(-20, xl + 20), rng.uniform(-20, yl + 20),
                              -rng.uniform(0.5, 40))
    return m, d


def oracle_kind(orc, kind):
    return orc.MAGGRAD if kind == nat.KIND_MAGNETIC_GRADIENT else orc.MAGNETIZATION


def ncols(kind, m):
    return 3 * m.ncells if kind == nat.KIND_MAGNETIZATION else m.ncells


def read_J(ctx):
    """J (rows x columns) as J * I in FP32, which is exact."""
    N, R = ctx.storedj_cols(), ctx.storedj_rows()
    eye = ctx.to_device(np.eye(N, dtype=np.float32))
    Y = ctx.alloc((N, R), np.float32)
    ctx.storedj_apply(nat.GEMM_FP32, eye, Y)
    return Y.to_host().T.astype(np.float64)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("host_block_cells", [None, 256])
def test_stored_J_matches_oracle(orc, kind, host_block_cells):
    m, d = make_model(kind, 3)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    ctx.storedj_build(kind, nat.FP64, host_block_cells)
    assert ctx.storedj_host_resident() == bool(host_block_cells)
    k = 3 if kind == nat.KIND_MAGNETIZATION else 1
    assert ctx.storedj_rows() == k * d.nsites
    assert ctx.storedj_cols() == ncols(kind, m)
    J = read_J(ctx)
    _, ref = orc.calculate(oracle_kind(orc, kind), m, m.GetModelParameters(), d.positions,
                           field=FIELD, want_sens=True)
    if kind == nat.KIND_MAGNETIC_GRADIENT:
        assert np.all(ref[:, m.ncells:] == 0.0)  # no background columns in the imp
    ref = ref[:, :ncols(kind, m)]
    assert J.shape == ref.shape
    tol = 2.0 ** -24 * np.abs(ref) + 1e-11 * np.max(np.abs(ref))
    assert np.all(np.abs(J - ref) <= tol * 1.0001)


@pytest.mark.parametrize("kind", KINDS)
def test_products_equal_fp64_path(orc, kind):
    m, d = make_model(kind, 5)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    ctx.storedj_build(kind, nat.FP64)
    mv = m.GetModelParameters()
    _, Jref = orc.calculate(oracle_kind(orc, kind), m, mv, d.positions, field=FIELD,
                            want_sens=True)
    Ja = np.abs(Jref[:, :ncols(kind, m)])
    n = ncols(kind, m)
    got = ctx.storedj_forward(nat.GEMM_FP32, mv)
    ref = ctx.forward(kind, nat.FP64, mv)
    bound = (n + 2) * 2.0 ** -24 * (Ja @ np.abs(mv[:n])) + 1e-11 * np.max(np.abs(ref))
    assert np.all(np.abs(got - ref) <= bound)
    r = np.random.default_rng(6).normal(size=ref.size)
    g = ctx.storedj_gradient(nat.GEMM_FP32, mv, r)
    gref = ctx.adjoint(kind, nat.FP64, mv, r)
    assert g.shape == gref.shape == (ctx.model_size(kind),)
    if kind == nat.KIND_MAGNETIC_GRADIENT:
        assert np.all(g[n:] == 0.0)
    gbound = 2 * (r.size + 2) * 2.0 ** -24 * (Ja.T @ np.abs(r)) + 1e-10 * np.max(np.abs(gref))
    assert np.all(np.abs(g[:n] - gref[:n]) <= gbound)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("calc_cls", [FullSensitivityGravMagCalculator, DiskGravMagCalculator])
def test_calculators(kind, calc_cls):
    m, d = make_model(kind, 8)
    if kind == nat.KIND_MAGNETIZATION:
        imp, direct_imp = HipMagnetizationImp(), HipMagnetizationImp()
    else:
        imp, direct_imp = HipMagneticGradImp(*FIELD), HipMagneticGradImp(*FIELD)
    calc = calc_cls(imp) if calc_cls is FullSensitivityGravMagCalculator else calc_cls(
        imp, block_cells=256)
    direct = MinMemGravMagCalculator(direct_imp)
    if kind == nat.KIND_MAGNETIZATION:
        calc.SetDataTransform(TotalField())
        direct.SetDataTransform(TotalField())
    first = calc.Calculate(m, d)  # builds J, direct result
    ref = direct.Calculate(m, d)
    assert np.array_equal(first, ref)
    rng = np.random.default_rng(9)
    mv = m.GetModelParameters()
    trials = [mv * (1 + 0.01 * rng.standard_normal(mv.size)) for _ in range(3)]
    cached = []
    for t in trials:
        if kind == nat.KIND_MAGNETIZATION:
            n = m.ncells
            m.SetMagnetization(t[:n], t[n:2 * n], t[2 * n:])
        else:
            m.SetSusceptibilities(t[:m.ncells])
            m.SetBackgroundSusceptibilities(t[m.ncells:])
        got = calc.Calculate(m, d)
        want = direct.Calculate(m, d)
        assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
        r = rng.normal(size=got.size)
        g = calc.LQDerivative(m, d, r)
        gw = direct.LQDerivative(m, d, r)
        assert np.max(np.abs(g - gw)) <= 1e-5 * np.max(np.abs(gw))
        cached.append((got, r, g))
    M = np.array(trials)
    R = np.array([c[1] for c in cached])
    F = calc.CalculateBatch(m, d, M)
    G = calc.LQDerivativeBatch(m, d, M, R)
    for b in range(len(trials)):
        assert np.array_equal(F[b], cached[b][0])
        assert np.array_equal(G[b], cached[b][2])


@pytest.mark.parametrize("kind", KINDS)
def test_sensitivity_rows_match_oracle(orc, kind):
    """FP64 rows in the reference layout: 1 row per site of ncells + nbg columns (zero
    background columns) or 3 rows of 3 ncells columns, against the oracle's matrix."""
    m, d = make_model(kind, 12)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    k = 3 if kind == nat.KIND_MAGNETIZATION else 1
    n = ctx.model_size(kind)
    s0, ns = 4, 9
    rows = ctx.alloc((ns * k, n))
    ctx.sensitivity_rows(kind, s0, ns, rows)
    ctx.synchronize()
    got = rows.to_host()
    _, ref = orc.calculate(oracle_kind(orc, kind), m, m.GetModelParameters(), d.positions,
                           field=FIELD, want_sens=True)
    ref = ref[s0 * k:(s0 + ns) * k]
    assert got.shape == ref.shape
    if kind == nat.KIND_MAGNETIC_GRADIENT:
        assert np.all(got[:, m.ncells:] == 0.0)
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))
    # the rows reproduce the forward: J m
    mv = m.GetModelParameters()
    fwd = ctx.forward(kind, nat.FP64, mv)[s0 * k:(s0 + ns) * k]
    assert np.max(np.abs(got @ mv - fwd)) <= 1e-11 * np.max(np.abs(fwd))


def test_field_change_drops_the_vertical_gradient_J():
    """The vertical-gradient J holds the inducing field's direction and strength
    (B = F/(4 pi G) T b, scaled by 1e-7/G): after another field is set the stored J is
    gone (NOT_READY), and a rebuilt J gives the direct forward of the new field."""
    kind = nat.KIND_MAGNETIC_GRADIENT
    m, d = make_model(kind, 21)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    ctx.storedj_build(kind, nat.FP64)
    assert ctx.storedj_rows() == d.nsites
    mv = m.GetModelParameters()
    before = ctx.storedj_forward(nat.GEMM_FP32, mv)
    ctx.set_field(FIELD[0], FIELD[1], FIELD[2])  # unchanged field keeps J
    assert ctx.storedj_rows() == d.nsites
    new_field = (0.3, 0.9, 47000.0)
    ctx.set_field(*new_field)
    assert ctx.storedj_rows() == 0
    out = ctx.alloc(d.nsites)
    status = ctx.lib.jgm_storedj_forward(ctx.handle, nat.GEMM_FP32,
                                         nat._devptr(ctx.to_device(mv)), nat._devptr(out))
    assert status == nat.JGM_ERR_NOT_READY
    ctx.storedj_build(kind, nat.FP64)
    got = ctx.storedj_forward(nat.GEMM_FP32, mv)
    want = ctx.forward(kind, nat.FP64, mv)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
    assert np.max(np.abs(got - before)) > 1e-3 * np.max(np.abs(want))
