"""Depth-weighting profile below the middle of the mesh from the GPU's adjoint
(CalculateMiddleSens, jif3D/Gravity/DepthWeighting.h:109-133), against the oracle's
LQDerivative for the same site and unit residual, in Python and through the C++ adapter."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from golden_paths import ROOT
from jif3d_amd.depth_weighting import (CalculateMiddleSens, FitZ0, ConstructDepthWeighting,
                                       WeightingTerm)
from jif3d_amd.imp import (ScalarHipGravityImp, TensorHipGravityImp, MinMemGravMagCalculator,
                           FullSensitivityGravMagCalculator)
from jif3d_amd.model import ThreeDGravityModel

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "jif3d_amd", "host", "adapter_driver")


def model(seed=3):
    rng = np.random.default_rng(seed)
    m = ThreeDGravityModel()
    m.SetMeshSize(8, 7, 12)
    m.SetXCellSizes(np.full(8, 25.0))
    m.SetYCellSizes(np.full(7, 30.0))
    m.SetZCellSizes(10.0 * 1.1 ** np.arange(12))
    m.SetDensities(rng.random(m.ncells) * 100 + 2000)
    m.SetBackgroundThicknesses([50.0, 1000.0])
    m.SetBackgroundDensities([2000.0, 2500.0])
    return m


@pytest.mark.parametrize("kind,calc", [(1, MinMemGravMagCalculator(ScalarHipGravityImp())),
                                       (2, MinMemGravMagCalculator(TensorHipGravityImp())),
                                       (1, FullSensitivityGravMagCalculator(ScalarHipGravityImp()))])
def test_middle_profile_matches_oracle(orc, kind, calc):
    m = model()
    prof = CalculateMiddleSens(m, calc)
    site = ([m.coords[0][-1] / 2.0], [m.coords[1][-1] / 2.0], [-0.01])
    k = 1 if kind == 1 else 9
    g = orc.lqderivative(kind, m, m.GetModelParameters(), site, np.ones(k))
    i, j = int(np.searchsorted(m.coords[0], site[0][0]) - 1), int(np.searchsorted(m.coords[1], site[1][0]) - 1)
    start = m.IndexToOffset(i, j, 0)
    ref = g[start:start + m.nz]
    # the stored-J calculator answers the gradient from its FP32 J
    tol = 2e-6 if isinstance(calc, FullSensitivityGravMagCalculator) else 1e-10
    assert np.max(np.abs(prof - ref)) <= tol * np.max(np.abs(ref))
    if kind == 1:  # gz sensitivity decays with depth; the fitted weights grow with it
        z0 = FitZ0(prof, m.GetZCellSizes(), WeightingTerm(-2.0))
        w = ConstructDepthWeighting(m.GetZCellSizes(), z0, WeightingTerm(-2.0))
        assert z0 > 0 and np.all(np.diff(w) > 0) and w[-1] == 1.0


def test_cpp_middle_sens_matches_python():
    from test_gpu_cpp_adapter import write_problem
    m = model(4)
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_problem(inp, 1, 0, m, ([1.0], [1.0], [-1.0]), (0.0, 0.0, 0.0), np.zeros(1))
        res = subprocess.run([DRIVER, inp, out, "--middle-sens"], capture_output=True, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stderr
        data = np.fromfile(out, dtype="<f8")
    nz = m.nz
    prof, z0, w = data[:nz], data[nz], data[nz + 1:]
    ref = CalculateMiddleSens(m, MinMemGravMagCalculator(ScalarHipGravityImp()))
    assert np.max(np.abs(prof - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert z0 == pytest.approx(FitZ0(ref, m.GetZCellSizes(), WeightingTerm(-2.0)), rel=1e-9)
    assert np.allclose(w, ConstructDepthWeighting(m.GetZCellSizes(), z0, WeightingTerm(-2.0)),
                       rtol=1e-12)
