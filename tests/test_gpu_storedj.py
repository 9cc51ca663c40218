"""Stored-J path on the GPU: the FP32 sensitivity matrix in HBM and its batched
products on the matrix cores, through the C ABI.

What pins it:
* J itself: every entry against the oracle's sensitivity rows (the reference's
  FullSensitivity row layout), to FP32 rounding of the FP64 value.  J is read back
  exactly as J * I (products with 1 and 0 and sums of zeros are exact in FP32).
* Products: against float64 products of the same FP32 J, within the a-priori error
  bound of each precision, err <= eps_eff * (|J| |M|) entrywise, with u = 2^-8 the
  unit roundoff of BF16 (8 significant bits) and K the length of one FP32 sum:
    FP32          (K + 2) 2^-24
    BF16          2u + u^2 + (K + 2) 2^-24        (both operands rounded to BF16)
    BF16 refined  3u^2 + 4u^3 + (K + 2) 2^-24     (x = hi + lo + e with |lo| <= u|x| and
                  |e| <= u^2 |x|; hi*hi + hi*lo + lo*hi drops lo*lo and the e terms)
  BF16 products are exact in the FP32 accumulator.
* Consistency with the matrix-free path: J m equals the FP64 forward (no background
  layers) and 2 J^T r the FP64 adjoint, to FP32 accuracy.
* Full batched size (64x64x32 cells, 32x32 sites, scalar+FTG, 64 right-hand sides):
  the FP64 forward of two trial models and the dot-product identity
  <J M, R> = <M, J^T R>."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd import synthetic
from jif3d_amd.model import ThreeDGravityModel, TensorGravityData

pytestmark = pytest.mark.gpu

TENSOR6 = [0, 1, 2, 4, 5, 8]  # Uxx Uxy Uxz Uyy Uyz Uzz out of the 9 raw tensor values
KROWS = {nat.KIND_SCALAR: 1, nat.KIND_TENSOR: 6, nat.KIND_MAGNETIC: 3,
         nat.KIND_SCALAR_TENSOR: 7}
PRECS = [nat.GEMM_FP32, nat.GEMM_BF16, nat.GEMM_BF16_REFINED]


def eps_eff(prec, K):
    acc = (K + 2) * 2.0 ** -24
    return {nat.GEMM_FP32: acc, nat.GEMM_BF16: 2.0 ** -7 + 2.0 ** -16 + acc,
            nat.GEMM_BF16_REFINED: 3 * 2.0 ** -16 + 2.0 ** -21 + acc}[prec]


def small_mesh(nx=12, ny=10, nz=8, nsx=7, nsy=6, bg=False, seed=5):
    rng = np.random.default_rng(seed)
    m = ThreeDGravityModel()
    m.SetMeshSize(nx, ny, nz)
    m.SetXCellSizes(rng.integers(10, 30, nx) + 0.0)
    m.SetYCellSizes(rng.integers(10, 30, ny) + 0.0)
    m.SetZCellSizes(rng.integers(5, 20, nz) + 0.0)
    m.SetDensities(2000.0 + 1000.0 * rng.random(nx * ny * nz))
    if bg:
        m.SetBackgroundThicknesses([50.0, 1e4])
        m.SetBackgroundDensities([2500.0, 2700.0])
    d = TensorGravityData()
    xl, yl = sum(m.sizes[0]), sum(m.sizes[1])
    for i in range(nsx):
        for j in range(nsy):
            d.AddMeasurementPoint((i + 0.37) * xl / nsx, (j + 0.61) * yl / nsy,
                                  -1.0 - 3.0 * ((i + j) % 4))
    return m, d


def oracle_compact_J(orc, kind, m, d, field=None, sites=None):
    """Compact rows (site-major, components fastest) x ncells from the oracle."""
    N = m.ncells
    sites = d.positions if sites is None else sites
    if kind == nat.KIND_SCALAR:
        return orc.calculate(orc.SCALAR, m, m.GetModelParameters(), sites, want_sens=True)[1][:, :N]
    if kind == nat.KIND_MAGNETIC:
        _, J = orc.calculate(orc.MAGNETIC, m, m.GetModelParameters(), sites, field=field,
                             want_sens=True)
        return J[:, :N]
    _, Jt = orc.calculate(orc.TENSOR, m, m.GetModelParameters(), sites, want_sens=True)
    Jt = Jt[:, :N].reshape(-1, 9, N)[:, TENSOR6, :]
    if kind == nat.KIND_TENSOR:
        return Jt.reshape(-1, N)
    _, Jg = orc.calculate(orc.SCALAR, m, m.GetModelParameters(), sites, want_sens=True)
    return np.concatenate([Jg[:, None, :N], Jt], axis=1).reshape(-1, N)


def make_ctx(m, d, field=None):
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    return ctx


def read_J(ctx):
    """J (rows x ncells) as J * I in FP32, which is exact."""
    N, R = ctx.storedj_cols(), ctx.storedj_rows()
    eye = ctx.to_device(np.eye(N, dtype=np.float32))
    Y = ctx.alloc((N, R), np.float32)
    ctx.storedj_apply(nat.GEMM_FP32, eye, Y)
    return Y.to_host().T.astype(np.float64)


def products(ctx, prec, M, Rr):
    R, N = ctx.storedj_rows(), ctx.storedj_cols()
    Md = ctx.to_device(M.astype(np.float32))
    Rd = ctx.to_device(Rr.astype(np.float32))
    Y = ctx.alloc((M.shape[0], R), np.float32)
    G = ctx.alloc((Rr.shape[0], N), np.float32)
    ctx.storedj_apply(prec, Md, Y)
    ctx.storedj_apply_t(prec, Rd, G)
    return Y.to_host().astype(np.float64), G.to_host().astype(np.float64)


@pytest.mark.parametrize("kind", [nat.KIND_SCALAR, nat.KIND_TENSOR, nat.KIND_SCALAR_TENSOR,
                                  nat.KIND_MAGNETIC])
def test_stored_J_matches_oracle_rows(orc, kind):
    m, d = small_mesh()
    field = (60.0, 10.0, 50000.0) if kind == nat.KIND_MAGNETIC else None
    ctx = make_ctx(m, d, field)
    ctx.storedj_build(kind, nat.FP64)
    k = KROWS[kind]
    assert ctx.storedj_rows() == k * d.nsites
    J = read_J(ctx)
    ref = oracle_compact_J(orc, kind, m, d, field)
    assert J.shape == ref.shape
    # FP32 rounding of each entry, plus the FP64 path's own 1e-11 of each component's scale
    scale = np.abs(ref.reshape(-1, k, ref.shape[1])).max(axis=(0, 2))
    tol = 2.0 ** -24 * np.abs(ref) + 1e-11 * np.repeat(scale[None, :], d.nsites, 0).reshape(-1, 1)
    assert np.all(np.abs(J - ref) <= tol * 1.0001)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("nrhs", [1, 3, 65])
def test_products_within_error_bound(prec, nrhs):
    """Ragged sizes: 960 cells, 42 sites x 7 rows, 1/3/65 right-hand sides."""
    m, d = small_mesh()
    ctx = make_ctx(m, d)
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64)
    J = read_J(ctx)
    R, N = J.shape
    rng = np.random.default_rng(nrhs)
    M = rng.normal(size=(nrhs, N)).astype(np.float32).astype(np.float64)
    Rr = rng.normal(size=(nrhs, R)).astype(np.float32).astype(np.float64)
    Y, G = products(ctx, prec, M, Rr)
    Yr, Gr = M @ J.T, Rr @ J
    bY = np.abs(M) @ np.abs(J).T
    bG = np.abs(Rr) @ np.abs(J)
    assert np.all(np.abs(Y - Yr) <= eps_eff(prec, N) * bY + 1e-45)
    assert np.all(np.abs(G - Gr) <= eps_eff(prec, R) * bG + 1e-45)
    if prec != nat.GEMM_FP32:
        # the BF16 paths do round their operands
        assert np.max(np.abs(Y - Yr)) > 0


def test_refined_is_close_to_fp32():
    m, d = small_mesh()
    ctx = make_ctx(m, d)
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64)
    N = ctx.storedj_cols()
    M = np.random.default_rng(0).normal(size=(8, N))
    Rr = np.random.default_rng(1).normal(size=(8, ctx.storedj_rows()))
    Y32, G32 = products(ctx, nat.GEMM_FP32, M, Rr)
    Yb, Gb = products(ctx, nat.GEMM_BF16, M, Rr)
    Yr, Gr = products(ctx, nat.GEMM_BF16_REFINED, M, Rr)
    eb = np.max(np.abs(Yb - Y32)) / np.max(np.abs(Y32))
    er = np.max(np.abs(Yr - Y32)) / np.max(np.abs(Y32))
    assert er < eb / 20
    eb = np.max(np.abs(Gb - G32)) / np.max(np.abs(G32))
    er = np.max(np.abs(Gr - G32)) / np.max(np.abs(G32))
    assert er < eb / 20


@pytest.mark.parametrize("kind", [nat.KIND_SCALAR_TENSOR, nat.KIND_MAGNETIC])
def test_J_times_model_equals_forward(kind):
    """J m (FP32) against the FP64 matrix-free forward without background layers."""
    m, d = small_mesh(bg=False)
    field = (60.0, 10.0, 50000.0) if kind == nat.KIND_MAGNETIC else None
    if kind == nat.KIND_MAGNETIC:
        m.SetDensities(1e-3 * (1 + np.random.default_rng(4).random(m.ncells)))
    ctx = make_ctx(m, d, field)
    ctx.storedj_build(kind, nat.MIXED)
    mv = m.GetModelParameters()
    f = ctx.forward(kind, nat.FP64, mv)
    if kind == nat.KIND_SCALAR_TENSOR:
        f = f.reshape(-1, 10)[:, [0] + [1 + i for i in TENSOR6]].reshape(-1)
    Y, _ = products(ctx, nat.GEMM_FP32, mv[None, :], np.zeros((1, ctx.storedj_rows())))
    k = KROWS[kind]
    fk, Yk = f.reshape(-1, k), Y.reshape(-1, k)
    for c in range(k):
        assert np.max(np.abs(Yk[:, c] - fk[:, c])) <= 2e-6 * np.max(np.abs(fk[:, c]))


def test_twice_JT_r_equals_adjoint():
    m, d = small_mesh(bg=False)
    ctx = make_ctx(m, d)
    ctx.storedj_build(nat.KIND_SCALAR, nat.FP64)
    r = np.random.default_rng(9).normal(size=d.nsites)
    g = ctx.adjoint(nat.KIND_SCALAR, nat.FP64, m.GetModelParameters(), r)
    _, G = products(ctx, nat.GEMM_FP32, np.zeros((1, m.ncells)), r[None, :])
    assert np.max(np.abs(2.0 * G[0] - g)) <= 2e-6 * np.max(np.abs(g))


def test_errors_and_release():
    m, d = small_mesh()
    ctx = make_ctx(m, d)
    N = m.ncells
    M = ctx.alloc((2, N), np.float32)
    status = ctx.lib.jgm_storedj_apply(ctx.handle, nat.GEMM_FP32, nat._devptr(M), 2,
                                       nat._devptr(M))
    assert status == nat.JGM_ERR_NOT_READY
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64)
    R = ctx.storedj_rows()
    Y = ctx.alloc((2, R), np.float32)
    with pytest.raises(ValueError):
        ctx.storedj_apply(nat.GEMM_FP32, M, ctx.alloc((3, R), np.float32))
    with pytest.raises(ValueError):
        ctx.storedj_apply(nat.GEMM_FP32, ctx.alloc((2, N)), Y)
    status = ctx.lib.jgm_storedj_apply(ctx.handle, nat.GEMM_FP32, nat._devptr(M), 4097,
                                       nat._devptr(Y))
    assert status == nat.JGM_ERR_UNSUPPORTED
    status = ctx.lib.jgm_storedj_apply(ctx.handle, 7, nat._devptr(M), 2, nat._devptr(Y))
    assert status == nat.JGM_ERR_INVALID_ARGUMENT
    ctx.storedj_apply(nat.GEMM_FP32, M, Y)
    ctx.storedj_release()
    assert ctx.storedj_rows() == 0
    status = ctx.lib.jgm_storedj_apply(ctx.handle, nat.GEMM_FP32, nat._devptr(M), 2,
                                       nat._devptr(Y))
    assert status == nat.JGM_ERR_NOT_READY


def test_batched_full_size(orc):
    """64x64x32 cells, 32x32 sites, scalar+FTG (7168 x 131072 FP32 J, 3.76 GB), 64 trial
    models and 64 residual vectors.  Rows of eight sites are checked against the oracle's
    FP64 rows within each precision's bound (plus the FP32 rounding of J); residuals
    restricted to those rows check J^T R the same way."""
    m, d, M, Rr = synthetic.config3()
    m.SetBackgroundThicknesses([])
    m.SetBackgroundDensities([])
    ctx = make_ctx(m, d)
    ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64)
    R, N = ctx.storedj_rows(), ctx.storedj_cols()
    assert (R, N) == (7168, 131072)
    M32 = M.astype(np.float32).astype(np.float64)
    R32 = Rr.astype(np.float32).astype(np.float64)
    sample = np.array([0, 31, 345, 496, 527, 700, 992, 1023])
    x, y, z = d.positions
    Js = oracle_compact_J(orc, nat.KIND_SCALAR_TENSOR, m, d,
                          sites=(x[sample], y[sample], z[sample]))
    rows = (sample[:, None] * 7 + np.arange(7)[None, :]).reshape(-1)
    Rs = np.zeros_like(R32)
    Rs[:, rows] = R32[:, rows]
    Yref, bY = M32 @ Js.T, np.abs(M32) @ np.abs(Js).T
    Gref, bG = Rs[:, rows] @ Js, np.abs(Rs[:, rows]) @ np.abs(Js)
    for prec in PRECS:
        Y, G = products(ctx, prec, M32, Rs)
        eY = eps_eff(prec, N) + 2.0 ** -24
        eG = eps_eff(prec, R) + 2.0 ** -24
        assert np.all(np.abs(Y[:, rows] - Yref) <= eY * bY), prec
        assert np.all(np.abs(G - Gref) <= eG * bG), prec
    # trial models 0 and 63 against the FP64 matrix-free forward at every site: the FTG
    # components cancel (sum |J| m is tens of times |J m|), hence 1e-5 of each maximum
    Y, G = products(ctx, nat.GEMM_FP32, M32, R32)
    for b in (0, 63):
        f = ctx.forward(nat.KIND_SCALAR_TENSOR, nat.FP64, M32[b])
        f = f.reshape(-1, 10)[:, [0] + [1 + i for i in TENSOR6]]
        Yb = Y[b].reshape(-1, 7)
        for c in range(7):
            assert np.max(np.abs(Yb[:, c] - f[:, c])) <= 1e-5 * np.max(np.abs(f[:, c]))
    # <J M, R> = <M, J^T R> for every right-hand side
    lhs = np.sum(Y * R32, axis=1)
    rhs = np.sum(M32 * G, axis=1)
    scale = np.sum(np.abs(Y) * np.abs(R32), axis=1)
    assert np.all(np.abs(lhs - rhs) <= 1e-5 * scale)


def test_bf16_copy_follows_a_rebuilt_J():
    """The BF16 J M of an HBM-resident J reads a packed BF16 copy of J made on first use: a
    rebuild (here of another kind, with other rows) or a release must not leave the old copy
    in use, and the same J gives the same products again."""
    m, d = small_mesh()
    ctx = make_ctx(m, d)
    rng = np.random.default_rng(7)
    N = m.ncells
    M = rng.normal(size=(5, N)).astype(np.float32).astype(np.float64)
    Rr = rng.normal(size=(5, 7 * d.nsites)).astype(np.float32).astype(np.float64)
    ys = []
    for step, kind in enumerate((nat.KIND_SCALAR_TENSOR, nat.KIND_SCALAR,
                                 nat.KIND_SCALAR_TENSOR)):
        ctx.storedj_build(kind, nat.FP64)
        J = read_J(ctx)
        R = J.shape[0]
        Y, _ = products(ctx, nat.GEMM_BF16, M, Rr[:, :R])
        assert Y.shape == (5, R)
        bound = eps_eff(nat.GEMM_BF16, N) * (np.abs(M) @ np.abs(J).T) + 1e-45
        assert np.all(np.abs(Y - M @ J.T) <= bound)
        ys.append(Y)
        if step == 1:
            ctx.storedj_release()
    assert np.array_equal(ys[0], ys[2])


@pytest.mark.parametrize("shape", [(12, 10, 8, 4, 4), (20, 17, 13, 11, 9)])
def test_bf16_packed_copy_equals_rounding_in_registers(monkeypatch, shape):
    """The BF16 copies of J hold the BF16 roundings of the FP32 J (in 256-row tiles of
    64-cell blocks for J M, rows past the last one zero; eight rows of a cell per 16 bytes in
    32-cell tiles for J^T R) and the products
    that read them use the same k order and flush interval, so they equal, bit for bit, the
    products that round the FP32 J in registers (the copies switched off with
    JIF3D_AMD_BF16_COPY=0 when the context is created)."""
    nx, ny, nz, nsx, nsy = shape
    m, d = small_mesh(nx, ny, nz, nsx, nsy)
    rng = np.random.default_rng(17)
    M = rng.normal(size=(6, m.ncells)).astype(np.float32)
    outs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("JIF3D_AMD_BF16_COPY", flag)
        ctx = make_ctx(m, d)
        ctx.storedj_build(nat.KIND_SCALAR_TENSOR, nat.FP64)
        R = ctx.storedj_rows()
        Rr = rng.normal(size=(70, R)).astype(np.float32) if flag == "1" else Rr
        Y = ctx.alloc((6, R), np.float32)
        G = ctx.alloc((70, m.ncells), np.float32)
        ctx.profile_enable(True)
        ctx.storedj_apply(nat.GEMM_BF16, ctx.to_device(M), Y)
        ctx.storedj_apply_t(nat.GEMM_BF16, ctx.to_device(Rr), G)
        ctx.synchronize()
        names = set(ctx.profile())
        assert ("storedj_jm_bf16_packed" in names) == (flag == "1"), names
        assert ("storedj_jtr_bf16_packed" in names) == (flag == "1"), names
        outs.append((Y.to_host(), G.to_host()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
