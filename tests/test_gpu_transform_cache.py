"""With a data transform, an adjoint at the model of the preceding forward takes the raw
data from that forward instead of computing it again.  The result must equal a context
that never ran the forward, bit for bit, and any change (model values, transform angles,
precision, geometry) must fall back to recomputing."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.model import ThreeDGravityModel, TensorGravityData

pytestmark = pytest.mark.gpu
FIELD = (1.0, 0.3, 50000.0)


def setup(seed=0):
    rng = np.random.default_rng(seed)
    m = ThreeDGravityModel()
    m.SetMeshSize(10, 9, 6)
    m.SetXCellSizes(rng.integers(10, 30, 10) + 0.0)
    m.SetYCellSizes(rng.integers(10, 30, 9) + 0.0)
    m.SetZCellSizes(rng.integers(5, 20, 6) + 0.0)
    m.SetDensities(rng.random(m.ncells) * 0.05)
    d = TensorGravityData()
    xl, yl = sum(m.sizes[0]), sum(m.sizes[1])
    for _ in range(25):
        d.AddMeasurementPoint(rng.uniform(0, xl), rng.uniform(0, yl), -rng.uniform(1, 40))
    return m, d, rng


def fresh_adjoint(m, d, kind, prec, tf, angles, model, r):
    c = nat.Context(0)
    c.configure(m, d, FIELD)
    c.set_transform(tf, *angles)
    return c.adjoint(kind, prec, model, r)


@pytest.mark.parametrize("kind,tf", [(nat.KIND_MAGNETIC, nat.TF_TOTAL_FIELD_ANOMALY),
                                     (nat.KIND_MAGNETIC, nat.TF_TOTAL_FIELD),
                                     (nat.KIND_TENSOR, nat.TF_FTG_INVARIANT)])
@pytest.mark.parametrize("prec", [nat.FP64, nat.MIXED])
def test_adjoint_after_forward_equals_fresh(kind, tf, prec):
    m, d, rng = setup(1)
    angles = (0.9, 0.2)
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    ctx.set_transform(tf, *angles)
    mv = m.GetModelParameters()
    f = ctx.forward(kind, prec, mv)
    r = rng.normal(size=f.size)
    g = ctx.adjoint(kind, prec, mv, r)
    assert np.array_equal(g, fresh_adjoint(m, d, kind, prec, tf, angles, mv, r))
    # a different model: recomputed
    mv2 = mv * 1.5
    g2 = ctx.adjoint(kind, prec, mv2, r)
    assert np.array_equal(g2, fresh_adjoint(m, d, kind, prec, tf, angles, mv2, r))


def test_cache_invalidated_by_settings():
    m, d, rng = setup(2)
    kind, tf = nat.KIND_MAGNETIC, nat.TF_TOTAL_FIELD_ANOMALY
    ctx = nat.Context(0)
    ctx.configure(m, d, FIELD)
    ctx.set_transform(tf, 0.9, 0.2)
    mv = m.GetModelParameters()
    r = rng.normal(size=d.nsites)
    ctx.forward(kind, nat.FP64, mv)
    ctx.set_transform(tf, 0.5, -0.4)  # new angles after the forward
    assert np.array_equal(ctx.adjoint(kind, nat.FP64, mv, r),
                          fresh_adjoint(m, d, kind, nat.FP64, tf, (0.5, -0.4), mv, r))
    ctx.forward(kind, nat.MIXED, mv)  # other precision
    assert np.array_equal(ctx.adjoint(kind, nat.FP64, mv, r),
                          fresh_adjoint(m, d, kind, nat.FP64, tf, (0.5, -0.4), mv, r))
    ctx.forward(kind, nat.FP64, mv)
    ctx.set_field(1.2, 0.1, 48000.0)  # other inducing field
    c = nat.Context(0)
    c.configure(m, d, (1.2, 0.1, 48000.0))
    c.set_transform(tf, 0.5, -0.4)
    assert np.array_equal(ctx.adjoint(kind, nat.FP64, mv, r), c.adjoint(kind, nat.FP64, mv, r))
    # device arrays: the model array changed in place after the forward
    dm = ctx.to_device(mv)
    out = ctx.alloc(d.nsites)
    ctx.forward(kind, nat.FP64, dm, out)
    dm.upload(mv * 2.0)
    dr, dg = ctx.to_device(r), ctx.alloc(ctx.n_params())
    ctx.adjoint(kind, nat.FP64, dm, dr, dg)
    ctx.synchronize()
    assert np.array_equal(dg.to_host(), c.adjoint(kind, nat.FP64, mv * 2.0, r))
