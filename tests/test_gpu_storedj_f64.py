"""FP64 stored J (jgm_storedj_build_fp64) and its products on the FP64 matrix cores.

The reference keeps J as a ublas::matrix<double> (jif3D/GravMag/
FullSensitivityGravMagCalculator.h:40-41, :116-130) and its caching tests ask the cached
result to equal the direct one to float epsilon, in percent (BOOST_CHECK_CLOSE(a, b,
FLT_EPSILON): 1.19e-9 relative):
* scalar_caching_test, jif3D/Gravity/test_ScalarCalculation.cpp:248-273;
* lqderivative_test, :275-306 (LQDerivative == 2 J^T r, direct and cached);
* tensor caching, jif3D/Gravity/test_TensorCalculation.cpp:110-184.
Restated here through FullSensitivityGravMagCalculator(gemm_precision=GEMM_FP64), plus the
FP64 J itself against the oracle's rows and the raw products against FP64 matrix
products at the 7168 x 131072 size of the batched stored-J configuration."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd import synthetic
from jif3d_amd.imp import (ScalarHipGravityImp, TensorHipGravityImp, HipMagneticSusceptibilityImp,
                           FullSensitivityGravMagCalculator, MinMemGravMagCalculator, FTGInvariant,
                           TotalFieldAnomaly)
from jif3d_amd.model import (ThreeDGravityModel, ThreeDSusceptibilityModel, ScalarGravityData,
                             TensorGravityData, TotalFieldMagneticData)

pytestmark = pytest.mark.gpu
FEPS = np.finfo(np.float32).eps / 100.0  # BOOST_CHECK_CLOSE(.., float eps) in percent
FIELD = (0.8, 0.25, 50000.0)


def random_model(seed, model, data, maxcells=10, nmeas=10, background=True):
    """MakeRandomModel (jif3D/Gravity/test_common.h:31-67), seeded."""
    rng = np.random.default_rng(seed)
    dims = [rng.integers(10, 35, rng.integers(1, maxcells + 1)).astype(float) for _ in range(3)]
    model.SetMeshSize(*(len(v) for v in dims))
    model.SetXCellSizes(dims[0])
    model.SetYCellSizes(dims[1])
    model.SetZCellSizes(dims[2])
    vals = rng.random(model.ncells) * 10 + 1.0
    bg = []
    if background:
        model.SetBackgroundThicknesses([50.0, 300.0])
        bg = [2.0, 3.0]
    model.SetModelParameters(np.concatenate([vals, bg]))
    xl, yl = int(np.floor(dims[0].sum())), int(np.floor(dims[1].sum()))
    for _ in range(nmeas):
        data.AddMeasurementPoint(rng.integers(0, xl) + 1.0, rng.integers(0, yl) + 1.0,
                                 -(rng.integers(0, 100) + 1.0))
    return model, data


def close_pct(a, b, tol=FEPS):
    """BOOST_CHECK_CLOSE: |a - b| within tol of each (entries that cancel to zero
    within tol of the largest)."""
    scale = np.maximum(np.minimum(np.abs(a), np.abs(b)), 1e-6 * np.max(np.abs(b)))
    return np.all(np.abs(a - b) <= tol * scale)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_scalar_caching(seed):
    m, d = random_model(seed, ThreeDGravityModel(), ScalarGravityData())
    calc = FullSensitivityGravMagCalculator(ScalarHipGravityImp(), gemm_precision=nat.GEMM_FP64)
    first = calc.Calculate(m, d)   # direct, builds the FP64 J
    second = calc.Calculate(m, d)  # J m
    assert close_pct(first, second)


@pytest.mark.parametrize("seed", [4, 5])
def test_lqderivative_direct_and_cached(orc, seed):
    m, d = random_model(seed, ThreeDGravityModel(), ScalarGravityData(), background=False)
    calc = FullSensitivityGravMagCalculator(ScalarHipGravityImp(), gemm_precision=nat.GEMM_FP64)
    misfit = np.random.default_rng(seed).random(d.nsites)
    deriv = calc.LQDerivative(m, d, misfit)  # no cache yet: matrix-free
    calc.Calculate(m, d)
    _, J = orc.calculate(orc.SCALAR, m, m.GetModelParameters(), d.positions, want_sens=True)
    compare = 2.0 * J.T @ misfit
    deriv2 = calc.LQDerivative(m, d, misfit)  # through the FP64 J
    assert close_pct(deriv, compare)
    assert close_pct(deriv2, compare)


@pytest.mark.parametrize("tf", [None, "invariant"])
def test_tensor_caching(tf):
    m, d = random_model(7, ThreeDGravityModel(), TensorGravityData())
    calc = FullSensitivityGravMagCalculator(TensorHipGravityImp(), gemm_precision=nat.GEMM_FP64)
    direct = MinMemGravMagCalculator(TensorHipGravityImp())
    if tf:
        calc.SetDataTransform(FTGInvariant())
        direct.SetDataTransform(FTGInvariant())
    first = calc.Calculate(m, d)
    second = calc.Calculate(m, d)
    assert close_pct(first, second)
    r = np.random.default_rng(8).normal(size=first.size)
    g, gd = calc.LQDerivative(m, d, r), direct.LQDerivative(m, d, r)
    # cached (per-cell J entries) and direct (corner-first) gradients sum differently
    # rounded terms: compared against the largest entry
    assert np.max(np.abs(g - gd)) <= 1e-12 * np.max(np.abs(gd))


def test_magnetic_tfa_caching():
    m, d = random_model(9, ThreeDSusceptibilityModel(), TotalFieldMagneticData())
    mv = m.GetModelParameters()
    m.SetModelParameters(mv * 1e-3)
    calc = FullSensitivityGravMagCalculator(HipMagneticSusceptibilityImp(*FIELD),
                                            gemm_precision=nat.GEMM_FP64)
    calc.SetDataTransform(TotalFieldAnomaly(FIELD[0], FIELD[1]))
    first = calc.Calculate(m, d)
    second = calc.Calculate(m, d)
    assert close_pct(first, second)


def test_fp64_J_matches_oracle_rows(orc):
    """J read back as J I (FP64 products are exact there) against the oracle's FP64 rows
    in the compact layout (gz, then Uxx Uxy Uxz Uyy Uyz Uzz), to 1e-13 of the largest."""
    m, d = random_model(11, ThreeDGravityModel(), TensorGravityData(), background=False)
    ctx = nat.Context(0)
    ctx.configure(m, d)
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64, fp64=True)
    N, R = ctx.storedj_cols(), ctx.storedj_rows()
    assert R == 7 * d.nsites
    eye = ctx.to_device(np.eye(N))
    Y = ctx.alloc((N, R))
    ctx.storedj_apply_f64(eye, Y)
    J = Y.to_host().T
    x = d.positions
    _, g = orc.calculate(orc.SCALAR, m, m.GetModelParameters(), x, want_sens=True)
    _, t = orc.calculate(orc.TENSOR, m, m.GetModelParameters(), x, want_sens=True)
    ref = np.empty_like(J)
    t = t.reshape(d.nsites, 9, -1)
    ref[0::7] = g[:, :N]
    for q, c in enumerate((0, 1, 2, 4, 5, 8)):
        ref[1 + q::7] = t[:, c, :N]
    assert np.max(np.abs(J - ref)) <= 1e-13 * np.max(np.abs(ref))
    # J^T I gives J^T
    eyeR = ctx.to_device(np.eye(R))
    G = ctx.alloc((R, N))
    ctx.storedj_apply_t_f64(eyeR, G)
    assert np.array_equal(G.to_host(), J)


def test_precision_mismatch_is_rejected():
    m, d = random_model(12, ThreeDGravityModel(), ScalarGravityData())
    ctx = nat.Context(0)
    ctx.configure(m, d)
    ctx.storedj_build(nat.KIND_SCALAR, nat.FP64, fp64=True)
    n = ctx.storedj_cols()
    M = ctx.to_device(np.ones((2, n), np.float32))
    Y = ctx.alloc((2, ctx.storedj_rows()), np.float32)
    st = ctx.lib.jgm_storedj_apply(ctx.handle, nat.GEMM_FP32, nat._devptr(M), 2, nat._devptr(Y))
    assert st == nat.JGM_ERR_UNSUPPORTED
    with pytest.raises(nat.JgmError):
        ctx.storedj_forward(nat.GEMM_FP32, m.GetModelParameters())
    ctx.storedj_build(nat.KIND_SCALAR, nat.FP64)  # FP32 J
    with pytest.raises(nat.JgmError):
        ctx.storedj_forward(nat.GEMM_FP64, m.GetModelParameters())


def test_full_size_products():
    """64x64x32 cells, 32x32 sites, scalar + tensor: J 7168 x 131072 in FP64 (7.5 GB).
    J M for 64 trial models against the FP64 matrix-free forward on two of them, and
    <J M, R> = <M, J^T R> for every right-hand side."""
    m, d, M, Rr = synthetic.config3()
    m.SetBackgroundThicknesses([])
    m.SetBackgroundDensities([])
    ctx = nat.Context(0)
    ctx.configure(m, d)
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64, fp64=True)
    R, N = ctx.storedj_rows(), ctx.storedj_cols()
    assert (R, N) == (7168, 131072)
    Md, Rd = ctx.to_device(M), ctx.to_device(Rr)
    Y, G = ctx.alloc((64, R)), ctx.alloc((64, N))
    ctx.storedj_apply_f64(Md, Y)
    ctx.storedj_apply_t_f64(Rd, G)
    Yh, Gh = Y.to_host(), G.to_host()
    # the stored entries are per-cell 8-corner sums (as the reference's J) and the forward
    # sums per corner: both FP64, rounded differently; the per-cell form loses ~1e-8 of an
    # entry for far cells, and the FTG components cancel, hence 1e-8 of each maximum
    for b in (0, 63):
        f = ctx.forward(nat.KIND_SCALAR_TENSOR, nat.FP64, M[b]).reshape(-1, 10)
        f = f[:, [0, 1, 2, 3, 5, 6, 9]]
        yb = Yh[b].reshape(-1, 7)
        for c in range(7):
            assert np.max(np.abs(yb[:, c] - f[:, c])) <= 1e-8 * np.max(np.abs(f[:, c]))
    lhs = np.sum(Yh * Rr, axis=1)
    rhs = np.sum(M * Gh, axis=1)
    scale = np.sum(np.abs(Yh) * np.abs(Rr), axis=1)
    assert np.all(np.abs(lhs - rhs) <= 1e-12 * scale)


def test_background_change_uses_the_cached_J():
    """The reference's cache (GridBackgroundModelCache.h:17-85) also treats a change of the
    background densities as a new model and recomputes directly; here the key is geometry
    only, and a new background density goes through the cached J m plus the stored layer
    terms.  Both give the same data: the cached result after the change equals a direct
    calculation to float epsilon (the reference's caching tolerance)."""
    m, d = random_model(13, ThreeDGravityModel(), ScalarGravityData())
    calc = FullSensitivityGravMagCalculator(ScalarHipGravityImp(), gemm_precision=nat.GEMM_FP64)
    calc.Calculate(m, d)
    m.SetBackgroundDensities([2.5, 3.5])
    cached = calc.Calculate(m, d)
    direct = MinMemGravMagCalculator(ScalarHipGravityImp()).Calculate(m, d)
    assert close_pct(cached, direct)
    r = np.random.default_rng(14).normal(size=d.nsites)
    g_cached = calc.LQDerivative(m, d, r)
    g_direct = MinMemGravMagCalculator(ScalarHipGravityImp()).LQDerivative(m, d, r)
    assert np.max(np.abs(g_cached - g_direct)) <= 1e-12 * np.max(np.abs(g_direct))
