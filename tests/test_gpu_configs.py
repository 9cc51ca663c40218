"""The survey's benchmark configurations at full size on the GPU.

config 1: FTG on 64x64x32 cells (25 m, 10 m layers growing 5 %) with 32x32 sites at
          z = -50 m: FP64 against the oracle's long double restatement on a subset of
          sites, mixed against FP64 on every site.
config 2: gradient of sum |d(m) - d_obs|^2 (gz + FTG) on 100x100x40 cells with 50x50
          sites; d_obs from a seeded true model plus 1 % Gaussian noise, start model
          2670 kg/m^3.  Checked against central differences in data space on 200 random
          cells (h = 0.1 kg/m^3; the misfit is quadratic, so the difference quotient has
          no truncation error) and against the oracle's gradient for a subset of sites.
config 4: magnetic B and total-field anomaly on 128x128x64 cells with 64x64 sites,
          mixed against FP64 forward and adjoint at the headline tolerance 1e-6."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd import synthetic

pytestmark = pytest.mark.gpu
TENSOR = nat.KIND_TENSOR
ST = nat.KIND_SCALAR_TENSOR


def comp_err(a, b, k):
    a, b = a.reshape(-1, k), b.reshape(-1, k)
    floor = 1e-6 * np.max(np.abs(b))
    return max(np.max(np.abs(a[:, c] - b[:, c])) / max(np.max(np.abs(b[:, c])), floor)
               for c in range(k))


def ctx_for(m, d, field=None):
    ctx = nat.Context(0)
    ctx.configure(m, d, field)
    return ctx


def test_config1_ftg(orc):
    m, d = synthetic.config1()
    ctx = ctx_for(m, d)
    mv = m.GetModelParameters()
    f64 = ctx.forward(TENSOR, nat.FP64, mv)
    mix = ctx.forward(TENSOR, nat.MIXED, mv)
    assert comp_err(mix, f64, 9) <= 1e-6
    sub = np.random.default_rng(11).choice(d.nsites, 24, replace=False)
    x, y, z = d.positions
    ld = orc.calculate_ld(orc.TENSOR, m, mv, (x[sub], y[sub], z[sub]))
    got = f64.reshape(-1, 9)[sub].reshape(-1)
    assert comp_err(got, ld, 9) <= 1e-11


def config2_problem():
    m, d = synthetic.config2()
    rng = np.random.default_rng(21)
    true = m.GetModelParameters()
    ctx = ctx_for(m, d)
    d_true = ctx.forward(ST, nat.FP64, true)
    noise = rng.normal(size=d_true.size).reshape(-1, 10)
    scale = np.abs(d_true.reshape(-1, 10)).max(axis=0)  # 1 % of each component's size
    d_obs = d_true + (0.01 * noise * scale).reshape(-1)
    start = true.copy()
    start[:m.ncells] = 2670.0
    return m, d, ctx, start, d_obs


def test_config2_gradient_fd():
    m, d, ctx, start, d_obs = config2_problem()
    r = ctx.forward(ST, nat.FP64, start) - d_obs
    g = ctx.adjoint(ST, nat.FP64, start, r)          # gradient of sum r^2 = 2 J^T r
    gmix = ctx.adjoint(ST, nat.MIXED, start, r)
    cells = np.random.default_rng(22).choice(m.ncells, 200, replace=False)
    h = 0.1
    mp = ctx.to_device(start)
    out_p = ctx.alloc(d.nsites * 10)
    out_m = ctx.alloc(d.nsites * 10)
    fd = np.empty(cells.size)
    for i, c in enumerate(cells):
        plus, minus = start.copy(), start.copy()
        plus[c] += h
        minus[c] -= h
        mp.upload(plus)
        ctx.forward(ST, nat.FP64, mp, out_p)
        mp.upload(minus)
        ctx.forward(ST, nat.FP64, mp, out_m)
        dd = out_p.to_host() - out_m.to_host()
        fd[i] = 2.0 * np.dot(r, dd) / (2 * h)
    gmax = np.max(np.abs(g))
    assert np.max(np.abs(g[cells] - fd)) <= 1e-7 * gmax
    assert np.max(np.abs(gmix[cells] - fd)) <= 1e-5 * gmax
    assert np.max(np.abs(gmix - g)) <= 1e-6 * gmax


def test_config2_gradient_vs_oracle(orc):
    """Residuals on 80 sites only (zero elsewhere): the GPU's 2 J^T r against the
    oracle's LQDerivative of the scalar and tensor calculators run separately."""
    m, d, ctx, start, d_obs = config2_problem()
    r = ctx.forward(ST, nat.FP64, start) - d_obs
    sub = np.sort(np.random.default_rng(23).choice(d.nsites, 80, replace=False))
    rs = np.zeros_like(r).reshape(-1, 10)
    rs[sub] = r.reshape(-1, 10)[sub]
    g = ctx.adjoint(ST, nat.FP64, start, rs.reshape(-1))
    x, y, z = d.positions
    sites = (x[sub], y[sub], z[sub])
    ref = (orc.lqderivative(orc.SCALAR, m, start, sites, rs[sub, 0].copy())
           + orc.lqderivative(orc.TENSOR, m, start, sites, rs[sub, 1:].reshape(-1).copy()))
    assert np.max(np.abs(g - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_config4_magnetic_full_size():
    m, d, field = synthetic.config4_magnetic()
    ctx = ctx_for(m, d, field)
    mv = m.GetModelParameters()
    kind = nat.KIND_MAGNETIC
    b64 = ctx.forward(kind, nat.FP64, mv)
    bmx = ctx.forward(kind, nat.MIXED, mv)
    assert comp_err(bmx, b64, 3) <= 1e-6
    ctx.set_transform(nat.TF_TOTAL_FIELD_ANOMALY, field[0], field[1])
    t64 = ctx.forward(kind, nat.FP64, mv)
    tmx = ctx.forward(kind, nat.MIXED, mv)
    assert np.max(np.abs(tmx - t64)) <= 1e-6 * np.max(np.abs(t64))
    r = np.random.default_rng(31).normal(size=d.nsites)
    g64 = ctx.adjoint(kind, nat.FP64, mv, r)
    gmx = ctx.adjoint(kind, nat.MIXED, mv, r)
    assert np.max(np.abs(gmx - g64)) <= 1e-6 * np.max(np.abs(g64))
