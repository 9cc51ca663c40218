"""The C++ host adapter (jif3d_amd/host/hip_gravmag_imp.hpp) through its driver program:
Calculate and LQDerivative of scalar, tensor and magnetic imps against the oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from golden_paths import ROOT
from jif3d_amd.model import ThreeDGravityModel

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "jif3d_amd", "host", "adapter_driver")


def write_problem(path, kind, prec, m, sites, field, misfit):
    nx, ny, nz = m.nx, m.ny, m.nz
    nbg = len(m.bg_thicknesses)
    hdr = np.array([kind, prec, nx, ny, nz, nbg, len(sites[0])], dtype="<i4")
    parts = list(m.coords) + list(m.sizes) + [m.bg_thicknesses] + list(sites)
    parts += [np.asarray(field, float), m.GetModelParameters(), misfit]
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        for p in parts:
            f.write(np.ascontiguousarray(p, dtype="<f8").tobytes())


@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("kind,okind,k", [(1, 1, 1), (2, 2, 9), (3, 3, 3)])
def test_cpp_adapter_matches_oracle(orc, kind, okind, k, full):
    """Direct imp (FP64, rounding-level agreement) and, with --full, the cached
    calculator over the FP32 stored J (2e-6 of the largest value)."""
    assert os.path.exists(DRIVER), "build the library first (python -m jif3d_amd.build)"
    rng = np.random.default_rng(40 + kind)
    m = ThreeDGravityModel()
    m.SetMeshSize(7, 9, 5)
    m.SetXCellSizes(rng.integers(10, 35, 7).astype(float))
    m.SetYCellSizes(rng.integers(10, 35, 9).astype(float))
    m.SetZCellSizes(rng.integers(10, 35, 5).astype(float))
    m.SetDensities(rng.random(7 * 9 * 5) * 10 + 1)
    m.SetBackgroundThicknesses([30.0, 500.0])
    m.SetBackgroundDensities([2.0, 3.0])
    ns = 11
    sites = (rng.random(ns) * 150, rng.random(ns) * 200, -rng.random(ns) * 50 - 1)
    field = (0.6, 0.2, 48000.0)
    misfit = rng.normal(size=ns * k)
    with tempfile.TemporaryDirectory() as td:
        inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_problem(inp, kind, 0, m, sites, field, misfit)
        cmd = [DRIVER, inp, out] + (["--full"] if full else [])
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stderr
        data = np.fromfile(out, dtype="<f8")
    fwd, grad = data[: ns * k], data[ns * k:]
    fld = field if kind == 3 else None
    ref = orc.calculate(okind, m, m.GetModelParameters(), sites, field=fld)
    gref = orc.lqderivative(okind, m, m.GetModelParameters(), sites, misfit, field=fld)
    tf, tg = (2e-6, 2e-6) if full else (1e-11, 1e-10)
    assert np.max(np.abs(fwd - ref)) <= tf * np.max(np.abs(ref))
    assert np.max(np.abs(grad - gref)) <= tg * np.max(np.abs(gref))


def test_cpp_adapter_reports_errors():
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "bad.bin")
        open(inp, "wb").write(b"\x01\x00")
        res = subprocess.run([DRIVER, inp, os.path.join(td, "o")], capture_output=True,
                             text=True, timeout=60)
        assert res.returncode == 1 and "cannot read header" in res.stderr
