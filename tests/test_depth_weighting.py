"""Depth weighting (jif3D/Gravity/DepthWeighting.cpp) on the host: the reference's own
test cases (jif3D/Gravity/test_weighting.cpp:22-91), the middle-site extraction, and the
C++ header (jif3d_amd/host/depth_weighting.hpp) against the Python restatement."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from golden_paths import ROOT
from jif3d_amd.depth_weighting import (WeightingTerm, FitZ0, ConstructDepthWeighting,
                                       ExtractMiddleSens, FindAssociatedIndices)
from jif3d_amd.model import ThreeDGravityModel, ScalarGravityData


def test_weighting_derivative():
    """test_weighting.cpp:22-33: analytic derivative against a central difference."""
    z0, n, z = np.arccos(-1.0), -2.0, 0.4
    f = WeightingTerm(n)
    delta = float(np.finfo(np.float32).eps)  # a double, as in the reference
    fd = (f(z + 2 * delta, z0) - f(z, z0)) / (2 * delta)
    assert abs(f.deriv(z + delta, z0) - fd) <= 1e-3 * abs(fd)


def test_weighting_average():
    """test_weighting.cpp:34-56: analytic interval average against a Riemann sum."""
    f = WeightingTerm(-3.0)
    z0, z1, z2, eps = 1.0, 200.0, 201.0, 0.001
    steps = int((z2 - z1) / eps)
    avg = np.mean([f(z1 + i * eps, z0) for i in range(steps)])
    assert abs(f.average(z1, z2, z0) - avg) <= 1e-5 * avg


def fitz0_case():
    nz, z0, n = 20, 237.0, -3.0
    zs = np.array([nz * i for i in range(nz)], dtype=float)
    prof = WeightingTerm(n)(np.cumsum(zs), z0)
    return zs, prof, z0, n


def test_fitz0_recovers_z0():
    """test_weighting.cpp:57-91: the fit recovers z0 = 237 within 1 %."""
    zs, prof, z0, n = fitz0_case()
    found = FitZ0(prof, zs, WeightingTerm(n))
    assert abs(found - z0) <= 0.01 * z0
    # sign and scale of the profile do not matter (|.| and normalisation)
    assert FitZ0(-3.7 * prof, zs, WeightingTerm(n)) == pytest.approx(found, rel=1e-12)


def test_construct_depth_weighting():
    w = ConstructDepthWeighting([10.0, 20.0, 30.0], 50.0, WeightingTerm(-2.0))
    bottoms = np.array([10.0, 30.0, 60.0])
    ref = (bottoms + 50.0) ** 2
    assert np.allclose(w, ref / ref.max(), rtol=1e-15)


def test_extract_middle_sens():
    """The row of the site closest to the middle, the column of cells under it
    (the structure of test_weighting.cpp:93-133 with a synthetic matrix)."""
    m = ThreeDGravityModel()
    m.SetMeshSize(3, 3, 3)
    for setter in (m.SetXCellSizes, m.SetYCellSizes, m.SetZCellSizes):
        setter([10.0] * 3)
    d = ScalarGravityData()
    for p in (2, 15, 28):
        d.AddMeasurementPoint(p, p, -1)
    k = 9
    sens = np.arange(3 * k * 27, dtype=float).reshape(3 * k, 27)
    prof = ExtractMiddleSens(m, d, sens, k)
    start = m.IndexToOffset(1, 1, 0)
    assert np.array_equal(prof, sens[k, start:start + 3])
    assert FindAssociatedIndices(m, 15.0, 15.0, 0.0) == (1, 1, 0)
    # on a boundary the lower cell; below the mesh clamped to 0, past its end index n
    assert FindAssociatedIndices(m, 10.0, -5.0, 25.0) == (0, 0, 2)
    assert FindAssociatedIndices(m, 35.0, 0.0, 0.0) == (3, 0, 0)


def test_cpp_header_matches_python():
    """jif3d_amd/host/depth_weighting.hpp compiled with g++: same z0 and weights."""
    src = r'''
#include <cstdio>
#include "depth_weighting.hpp"
int main() {
  using namespace jif3d_amd;
  std::vector<double> zs(20), prof(20);
  double d = 0;
  for (int i = 0; i < 20; ++i) { zs[i] = 20.0 * i; d += zs[i]; prof[i] = WeightingTerm(-3.0)(d, 237.0); }
  std::printf("%.17g\n", FitZ0(prof, zs, WeightingTerm(-3.0)));
  for (double v : ConstructDepthWeighting({10, 20, 30}, 50.0, WeightingTerm(-2.0))) std::printf("%.17g\n", v);
  std::printf("%zu %zu\n", AssociatedIndex({0, 10, 20, 30}, 15.0), AssociatedIndex({0, 10, 20, 30}, -4.0));
}
'''
    with tempfile.TemporaryDirectory() as td:
        cpp, exe = os.path.join(td, "t.cpp"), os.path.join(td, "t")
        open(cpp, "w").write(src)
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "jif3d_amd", "host"),
                        "-I", os.path.join(ROOT, "include"), cpp, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    zs, prof, z0, n = fitz0_case()
    assert float(out[0]) == FitZ0(prof, zs, WeightingTerm(n))
    w = ConstructDepthWeighting([10.0, 20.0, 30.0], 50.0, WeightingTerm(-2.0))
    assert np.array_equal(np.array([float(v) for v in out[1:4]]), w)
    assert out[4].split() == ["1", "0"]
