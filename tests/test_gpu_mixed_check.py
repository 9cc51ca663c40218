"""The mixed path's a-posteriori check (jgm_set_mixed_check): after a mixed forward the
FP64 closed form is evaluated at a few sample sites (over every corner of the mesh), after a
mixed adjoint at a few sample cells (over every site), and when the sampled error exceeds
the tolerance, relative to each component's survey maximum (or to the largest gradient
entry), the call is recomputed in FP64.

The case it exists for: a nearly uniformly magnetised wide slab (susceptibility 1 +- 1 %
in the mesh, continued sideways by background layers of susceptibility 1), whose external
field is a small remainder of its cells' fields, so the far field's FP32 error per term is
amplified far beyond the stated 1e-6.  The headline-like gravity model passes the check and
keeps the mixed result."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd import synthetic
from jif3d_amd.model import ThreeDSusceptibilityModel, TotalFieldMagneticData

pytestmark = pytest.mark.gpu
FIELD = (1.05, 0.17, 50000.0)


def comp_err(a, b, k):
    a, b = a.reshape(-1, k), b.reshape(-1, k)
    return max(np.max(np.abs(a[:, c] - b[:, c])) / np.max(np.abs(b[:, c])) for c in range(k))


def slab():
    rng = np.random.default_rng(31)
    m = ThreeDSusceptibilityModel()
    n = (48, 44, 16)
    m.SetMeshSize(*n)
    for f, k in zip((m.SetXCellSizes, m.SetYCellSizes, m.SetZCellSizes), n):
        f(np.full(k, 10.0))
    m.SetSusceptibilities(1.0 + 0.01 * rng.uniform(-1.0, 1.0, m.ncells))
    m.SetBackgroundThicknesses([160.0, 1e4])
    m.SetBackgroundSusceptibilities([1.0, 0.0])
    d = TotalFieldMagneticData()
    for i in range(12):
        for j in range(12):
            d.AddMeasurementPoint((i + 0.3) * 40.0, (j + 0.6) * 36.7, -rng.uniform(1, 30))
    return m, d


def test_cancelling_slab_falls_back_to_fp64():
    m, d = slab()
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    mv = m.GetModelParameters()
    K = nat.KIND_MAGNETIC
    ref = ctx.forward(K, nat.FP64, mv)
    ctx.set_mixed_check(0, 0, 5e-7)  # off: the mixed result as it is
    unchecked = ctx.forward(K, nat.MIXED, mv)
    assert comp_err(unchecked, ref, 3) > 1e-6  # the bound does not hold on this model
    ctx.set_mixed_check(8, 16, 5e-7)
    _, _, before = ctx.mixed_check_report()
    checked = ctx.forward(K, nat.MIXED, mv)
    fe, _, after = ctx.mixed_check_report()
    assert fe > 5e-7 and after == before + 1
    assert comp_err(checked, ref, 3) <= 1e-12  # recomputed in FP64
    r = np.random.default_rng(32).normal(size=ref.size)
    g64 = ctx.adjoint(K, nat.FP64, mv, r)
    g = ctx.adjoint(K, nat.MIXED, mv, r)
    _, ae, _ = ctx.mixed_check_report()
    assert np.max(np.abs(g - g64)) <= 1e-6 * np.max(np.abs(g64))
    assert ae >= 0.0


def test_headline_like_gravity_keeps_the_mixed_result():
    m, d = synthetic.config4_gravity(n=48, nzc=24, nsite=24)
    ctx = nat.Context(0)
    ctx.configure(m, d)
    mv = m.GetModelParameters()
    K = nat.KIND_SCALAR_TENSOR
    ref = ctx.forward(K, nat.FP64, mv)
    ctx.set_mixed_check(8, 16, 5e-7)
    _, _, before = ctx.mixed_check_report()
    mix = ctx.forward(K, nat.MIXED, mv)
    fe, _, after = ctx.mixed_check_report()
    assert after == before  # no fallback
    assert 0.0 < fe <= 5e-7  # the sample saw the mixed path's own error
    assert 0.0 < comp_err(mix, ref, 10) <= 1e-6
    r = np.random.default_rng(33).normal(size=ref.size) * 1e-9
    g = ctx.adjoint(K, nat.MIXED, mv, r)
    _, ae, after2 = ctx.mixed_check_report()
    assert after2 == before and ae <= 5e-7
    g64 = ctx.adjoint(K, nat.FP64, mv, r)
    assert np.max(np.abs(g - g64)) <= 1e-6 * np.max(np.abs(g64))


def test_check_settings_are_validated():
    ctx = nat.Context(0)
    with pytest.raises(nat.JgmError):
        ctx.set_mixed_check(8, 16, -1.0)
    fe, ae, n = ctx.mixed_check_report()
    assert n == 0
