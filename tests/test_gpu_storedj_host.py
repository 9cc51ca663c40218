"""Stored J in page-locked host memory, streamed to HBM in column blocks (the role of
jif3D's DiskGravMagCalculator), against the same J in HBM: J^T R bit for bit, J M to the
FP32 rounding of a different grouping of block sums; through the C ABI and the Python
calculator."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.imp import (ScalarTensorHipGravityImp, DiskGravMagCalculator,
                           FullSensitivityGravMagCalculator)
from jif3d_amd.model import ThreeDGravityModel, TensorGravityData

pytestmark = pytest.mark.gpu


def problem(seed=0, nx=20, ny=15, nz=10, ns=30):
    rng = np.random.default_rng(seed)
    m = ThreeDGravityModel()
    m.SetMeshSize(nx, ny, nz)
    m.SetXCellSizes(rng.integers(10, 30, nx) + 0.0)
    m.SetYCellSizes(rng.integers(10, 30, ny) + 0.0)
    m.SetZCellSizes(rng.integers(5, 20, nz) + 0.0)
    m.SetDensities(1.0 + rng.random(m.ncells))
    m.SetBackgroundThicknesses([50.0, 500.0])
    m.SetBackgroundDensities([1.0, 2.0])
    d = TensorGravityData()
    xl, yl = sum(m.sizes[0]), sum(m.sizes[1])
    for _ in range(ns):
        d.AddMeasurementPoint(rng.uniform(0, xl), rng.uniform(0, yl), -rng.uniform(1, 40))
    return m, d, rng


def contexts(m, d, block):
    kind = nat.KIND_SCALAR_TENSOR
    a, b = nat.Context(0), nat.Context(0)
    for c in (a, b):
        c.configure(m, d)
    a.storedj_build(kind, nat.FP64)
    b.storedj_build(kind, nat.FP64, host_block_cells=block)
    assert not a.storedj_host_resident() and b.storedj_host_resident()
    return a, b


@pytest.mark.parametrize("gemm", [nat.GEMM_FP32, nat.GEMM_BF16, nat.GEMM_BF16_REFINED])
def test_products_match_hbm_resident(gemm):
    m, d, rng = problem()
    a, b = contexts(m, d, block=256)  # 3000 cells -> 12 blocks
    R, N = a.storedj_rows(), m.ncells
    nrhs = 5
    M = rng.normal(size=(nrhs, N)).astype(np.float32)
    Rr = rng.normal(size=(nrhs, R)).astype(np.float32)
    out = {}
    for name, c in (("a", a), ("b", b)):
        dM, dR = c.to_device(M), c.to_device(Rr)
        Y, G = c.alloc((nrhs, R), np.float32), c.alloc((nrhs, N), np.float32)
        c.storedj_apply(gemm, dM, Y)
        c.storedj_apply_t(gemm, dR, G)
        c.synchronize()
        out[name] = (Y.to_host(), G.to_host())
    Ya, Ga = out["a"]
    Yb, Gb = out["b"]
    assert np.array_equal(Ga, Gb)
    assert np.max(np.abs(Ya - Yb)) <= 1e-6 * np.max(np.abs(Ya))


def test_forward_gradient_batches_and_rebuild():
    m, d, rng = problem(1)
    a, b = contexts(m, d, block=512)
    kind = nat.KIND_SCALAR_TENSOR
    X = m.GetModelParameters()[None, :] * (1 + 0.01 * rng.random((4, m.ncells + 2)))
    Fa, Fb = a.storedj_forward_batch(nat.GEMM_FP32, X), b.storedj_forward_batch(nat.GEMM_FP32, X)
    assert np.max(np.abs(Fa - Fb)) <= 1e-6 * np.max(np.abs(Fa))
    Rr = rng.normal(size=Fa.shape)
    Ga = a.storedj_gradient_batch(nat.GEMM_FP32, None, Rr)
    Gb = b.storedj_gradient_batch(nat.GEMM_FP32, None, Rr)
    assert np.array_equal(Ga, Gb)
    # against the direct FP64 path: the FP32 J's accuracy
    D = b.forward_batch(kind, nat.FP64, X)
    assert np.max(np.abs(Fb - D)) <= 2e-6 * np.max(np.abs(D))
    # rebuilding in HBM and releasing leave a consistent context
    b.storedj_build(kind, nat.FP64)
    assert not b.storedj_host_resident()
    assert np.array_equal(b.storedj_gradient_batch(nat.GEMM_FP32, None, Rr), Ga)
    b.storedj_release()
    with pytest.raises(nat.JgmError):
        b.storedj_forward(nat.GEMM_FP32, X[0])


def test_disk_calculator_matches_full_sensitivity():
    m, d, rng = problem(2)
    full = FullSensitivityGravMagCalculator(ScalarTensorHipGravityImp())
    disk = DiskGravMagCalculator(ScalarTensorHipGravityImp(), block_cells=256)
    first_f, first_d = full.Calculate(m, d), disk.Calculate(m, d)
    assert np.array_equal(first_f, first_d)  # first call: the direct FP64 result
    m.SetDensities(m.GetDensities().ravel() * 1.1)
    cf, cd = full.Calculate(m, d), disk.Calculate(m, d)
    assert np.max(np.abs(cf - cd)) <= 1e-6 * np.max(np.abs(cf))
    r = rng.normal(size=cf.size)
    assert np.array_equal(full.LQDerivative(m, d, r), disk.LQDerivative(m, d, r))


@pytest.mark.parametrize("block", [256, 768, 100000])
def test_file_backed_J_equals_host_resident(tmp_path, block):
    """J in a file (the reference's Disk calculator keeps it in a file in a temporary
    directory, DiskGravMagCalculator.h:64-85): read back block by block, the products equal
    those of the host-resident J with the same blocks bit for bit; the file exists while J
    does and is removed on release."""
    m, d, rng = problem(7)
    kind = nat.KIND_SCALAR_TENSOR
    a, b = nat.Context(0), nat.Context(0)
    for c in (a, b):
        c.configure(m, d)
    a.storedj_build(kind, nat.FP64, host_block_cells=block)
    b.storedj_build(kind, nat.FP64, host_block_cells=block, directory=str(tmp_path))
    assert b.storedj_file_backed() and not b.storedj_host_resident()
    files = list(tmp_path.iterdir())
    assert len(files) == 1 and files[0].stat().st_size >= 4 * b.storedj_rows() * m.ncells
    R, N = a.storedj_rows(), a.storedj_cols()
    M = rng.normal(size=(5, N)).astype(np.float32)
    Rr = rng.normal(size=(5, R)).astype(np.float32)
    outs = []
    for c in (a, b):
        Y, G = c.alloc((5, R), np.float32), c.alloc((5, N), np.float32)
        c.storedj_apply(nat.GEMM_FP32, c.to_device(M), Y)
        c.storedj_apply_t(nat.GEMM_FP32, c.to_device(Rr), G)
        outs.append((Y.to_host(), G.to_host()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    mv = m.GetModelParameters()
    r = rng.normal(size=d.nsites * 10)
    assert np.array_equal(a.storedj_forward(nat.GEMM_FP32, mv), b.storedj_forward(nat.GEMM_FP32, mv))
    assert np.array_equal(a.storedj_gradient(nat.GEMM_FP32, mv, r),
                          b.storedj_gradient(nat.GEMM_FP32, mv, r))
    b.storedj_release()
    assert not list(tmp_path.iterdir())


def test_disk_calculator_with_a_directory(tmp_path):
    """DiskGravMagCalculator(imp, TempDir=...): direct first result, cached results through
    the file-backed J, against the direct imp."""
    m, d, rng = problem(8)
    calc = DiskGravMagCalculator(ScalarTensorHipGravityImp(), block_cells=512,
                                 TempDir=str(tmp_path))
    direct = ScalarTensorHipGravityImp()
    first = calc.Calculate(m, d)
    assert np.array_equal(first, direct.Calculate(m, d))
    assert len(list(tmp_path.iterdir())) == 1
    mv = m.GetModelParameters()
    m.SetModelParameters(mv * (1 + 0.01 * rng.standard_normal(mv.size)))
    got, want = calc.Calculate(m, d), direct.Calculate(m, d)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
    r = rng.normal(size=got.size)
    g, gw = calc.LQDerivative(m, d, r), direct.LQDerivative(m, d, r)
    assert np.max(np.abs(g - gw)) <= 1e-5 * np.max(np.abs(gw))


def test_file_backed_J_errors(tmp_path):
    """A directory that does not exist is reported (no J, no file left); bad block sizes
    are rejected; a later build in a good directory works."""
    m, d, rng = problem(9)
    kind = nat.KIND_SCALAR
    ctx = nat.Context(0)
    ctx.configure(m, d)
    with pytest.raises(nat.JgmError) as e:
        ctx.storedj_build(kind, nat.FP64, host_block_cells=256, directory=str(tmp_path / "nope"))
    assert "nope" in str(e.value)
    assert ctx.storedj_rows() == 0 and not ctx.storedj_file_backed()
    st = ctx.lib.jgm_storedj_build_file(ctx.handle, kind, nat.FP64, 0, str(tmp_path).encode())
    assert st == nat.JGM_ERR_INVALID_ARGUMENT
    assert not list(tmp_path.iterdir())
    ctx.storedj_build(kind, nat.FP64, host_block_cells=256, directory=str(tmp_path))
    assert ctx.storedj_file_backed() and ctx.storedj_rows() == d.nsites
    mv = m.GetModelParameters()
    want = ctx.forward(kind, nat.FP64, mv)
    got = ctx.storedj_forward(nat.GEMM_FP32, mv)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
