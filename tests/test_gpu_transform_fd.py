"""Transformed gradients on the GPU against finite differences of the GPU's own misfit
and against the oracle: the GPU side of the reference's transform-chain tests
(jif3D/Magnetics/test_MagneticObjective.cpp:99-134 T_deriv_test, :136-175 deltaT_deriv_test,
jif3D/Gravity/test_transforms.cpp:124-177 for the FTG invariant).

Setup as there: a random model of at most 5 cells per axis (MakeRandomModel,
jif3D/Gravity/test_common.h:31-67, seeded here), 3 sites, observed data 1.1 times the
synthetic data, unit errors; misfit sum (T(F(m)) - d_obs)^2, gradient LQDerivative(m, T(F(m))
- d_obs), central differences, tolerance 0.01 % (the reference's).  The reference steps by
1 % of each parameter; the total field's square root leaves an O(h^2) truncation of up to
3e-4 there on some seeds, above that tolerance, so the steps here are 0.1 %.  Both
precisions; the model is small, so the mixed path's near boxes hold every cell and it
equals the FP64 closed form."""
import numpy as np
import pytest

from jif3d_amd import _native as nat
from jif3d_amd.imp import (TensorHipGravityImp, HipMagneticSusceptibilityImp, FTGInvariant,
                           TotalField, TotalFieldAnomaly, MinMemGravMagCalculator)
from jif3d_amd.model import (ThreeDGravityModel, ThreeDSusceptibilityModel, TensorGravityData,
                             TotalFieldMagneticData)

pytestmark = pytest.mark.gpu


def random_model(rng, model, data, maxcells=5, nmeas=3):
    dims = [rng.integers(10, 35, rng.integers(1, maxcells + 1)).astype(float) for _ in range(3)]
    model.SetMeshSize(*(len(v) for v in dims))
    model.SetXCellSizes(dims[0])
    model.SetYCellSizes(dims[1])
    model.SetZCellSizes(dims[2])
    model.SetModelParameters(rng.random(model.ncells) * 10 + 1.0)
    xl, yl = int(np.floor(dims[0].sum())), int(np.floor(dims[1].sum()))
    for _ in range(nmeas):
        data.AddMeasurementPoint(rng.integers(0, xl) + 1.0, rng.integers(0, yl) + 1.0,
                                 -(rng.integers(0, 100) + 1.0))
    return model, data


def setup(tf_name, seed, prec):
    rng = np.random.default_rng(seed)
    if tf_name == "ftg_invariant":
        m, d = random_model(rng, ThreeDGravityModel(), TensorGravityData())
        imp = TensorHipGravityImp(prec)
        tf, okind, field, angles = FTGInvariant(), "TENSOR", None, (0.0, 0.0)
    else:
        m, d = random_model(rng, ThreeDSusceptibilityModel(), TotalFieldMagneticData())
        inc, dec, fs = rng.random(), rng.random(), 1.0 + rng.random()
        imp = HipMagneticSusceptibilityImp(inc, dec, fs, prec)
        tf = TotalField() if tf_name == "total_field" else TotalFieldAnomaly(inc, dec)
        okind, field, angles = "MAGNETIC", (inc, dec, fs), (inc, dec)
    calc = MinMemGravMagCalculator(imp)
    calc.SetDataTransform(tf)
    return m, d, calc, (okind, field, tf.code, angles)


@pytest.mark.parametrize("prec", [nat.FP64, nat.MIXED])
@pytest.mark.parametrize("tf_name", ["total_field", "total_field_anomaly", "ftg_invariant"])
@pytest.mark.parametrize("seed", [11, 12])
def test_transformed_gradient_fd(orc, tf_name, seed, prec):
    m, d, calc, (okind, field, tfc, angles) = setup(tf_name, seed, prec)
    mv = m.GetModelParameters()

    def fwd(x):
        m.SetModelParameters(x)
        return calc.Calculate(m, d)

    obs = 1.1 * fwd(mv)

    def misfit(x):
        r = fwd(x) - obs
        return float(r @ r)

    diff = fwd(mv) - obs
    grad = calc.LQDerivative(m, d, diff)
    assert grad.shape == mv.shape
    for i in range(mv.size):
        h = 0.001 * mv[i]
        xp, xm = mv.copy(), mv.copy()
        xp[i] += h
        xm[i] -= h
        fd = (misfit(xp) - misfit(xm)) / (2 * h)
        if abs(fd) > 1e-20:
            assert abs(fd - grad[i]) <= 1e-4 * max(abs(fd), abs(grad[i])), (i, fd, grad[i])
        else:
            assert abs(grad[i]) < 1e-10
    # and against the oracle's transformed LQDerivative at the same misfit
    m.SetModelParameters(mv)
    gref = orc.lqderivative(getattr(orc, okind), m, mv, d.positions, diff, field=field,
                            transform=tfc, tf_angles=angles)
    assert np.max(np.abs(grad - gref)) <= 1e-11 * np.max(np.abs(gref))
