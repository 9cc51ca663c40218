"""Depth weighting of the gravity and magnetic sensitivities
(jif3D/Gravity/DepthWeighting.h:24-133, DepthWeighting.cpp:14-160).

The sensitivity of a surface site decays with the depth of the cell; inversions scale the
model covariance by w(z) = 1 / (z + z0)^n, normalised to a maximum of 1, with z0 fitted
to the decay below a site in the middle of the mesh.  CalculateMiddleSens obtains that
decay from the GPU's matrix-free adjoint (one site, unit residual), as the reference
obtains it from LQDerivative (DepthWeighting.h:109-133); the rest is host arithmetic on
nz values.
"""
import numpy as np

from .model import GeneralData


class WeightingTerm:
    """f(z, z0) = (z + z0)^n with its z-derivative and interval average."""

    def __init__(self, exponent):
        self.n = float(exponent)

    def __call__(self, z, z0):
        return (z + z0) ** self.n

    def deriv(self, z, z0):
        return self.n * (z + z0) ** (self.n - 1.0)

    def average(self, z1, z2, z0):
        return 1.0 / ((z2 - z1) * (self.n + 1.0)) * ((z2 + z0) ** (self.n + 1.0)
                                                     - (z1 + z0) ** (self.n + 1.0))


def FitZ0(sens_profile, zsizes, weight_function):
    """z0 for which the normalised f(z_bottom_k, z0) matches the normalised |profile|:
    the reference's one-parameter Gauss-Newton iteration (DepthWeighting.cpp:52-116),
    starting at 1.1 times the mesh depth, at most 100 steps, until a step is below
    0.01 m.  (The reference also writes the fitted curve to a file "match.out"; this
    restatement has no side effects.)"""
    profile = np.abs(np.asarray(sens_profile, dtype=np.float64))
    zvalues = np.cumsum(np.asarray(zsizes, dtype=np.float64))
    profile = profile / np.max(profile)
    z0 = zvalues[-1] * 1.1
    stepsize, it = 1e6, 0
    while stepsize > 0.01 and it < 100:
        calculated = np.array([weight_function(z, z0) for z in zvalues])
        maximum = np.max(calculated)
        calculated = calculated / maximum
        sens = np.array([weight_function.deriv(z, z0) for z in zvalues]) / maximum
        delta = profile - calculated
        # ModelSpaceInversion with one parameter, unit errors and weights, lambda = 0
        # (jif3D/Inversion/LinearInversion.cpp:109-147): (s^T s)^-1 s^T delta
        step = np.dot(sens, delta) / np.dot(sens, sens)
        stepsize = abs(step)
        z0 -= step
        it += 1
    return z0


def ConstructDepthWeighting(zsizes, z0, weight_function):
    """1 / f(bottom of layer k, z0) for each of the nz layers, normalised to max 1
    (DepthWeighting.cpp:128-152)."""
    bottoms = np.cumsum(np.asarray(zsizes, dtype=np.float64))
    w = np.array([1.0 / weight_function(z, z0) for z in bottoms])
    return w / np.max(w)


def FindAssociatedIndices(model, x, y, z):
    """Cell indices containing (x, y, z), clamped at 0 (ThreeDModelBase.cpp:137-156)."""
    out = []
    for coords, v in zip(model.coords, (x, y, z)):
        idx = int(np.searchsorted(np.asarray(coords), v, side="left"))
        out.append(max(idx - 1, 0))
    return tuple(out)


def ExtractMiddleSens(model, data, sensitivities, meas_per_pos):
    """The nz sensitivities below the site closest to the middle (last x and y
    coordinate halved, as DepthWeighting.h:62-103) from a row-major sensitivity matrix
    with meas_per_pos rows per site."""
    xs, ys = data.GetMeasPosX(), data.GetMeasPosY()
    midx = model.coords[0][-1] / 2.0
    midy = model.coords[1][-1] / 2.0
    mid = int(np.argmin(np.sqrt((xs - midx) ** 2 + (ys - midy) ** 2)))
    i, j, _ = FindAssociatedIndices(model, xs[mid], ys[mid], 0.0)
    row = np.asarray(sensitivities)[mid * meas_per_pos]
    start = (model.nz * model.ny) * i + model.nz * j
    return np.array(row[start:start + model.nz])


def CalculateMiddleSens(model, calculator):
    """Sensitivity decay below a site at the middle of the mesh, 1 cm above its top:
    Calculate, then LQDerivative with a unit residual for every datum of the site
    (DepthWeighting.h:109-133), on the GPU through `calculator`."""
    data = GeneralData()
    data.AddMeasurementPoint(model.coords[0][-1] / 2.0, model.coords[1][-1] / 2.0, -0.01)
    calculator.Calculate(model, data)
    misfit = np.ones(calculator.GetDataPerMeasurement())
    deriv = calculator.LQDerivative(model, data, misfit)
    return ExtractMiddleSens(model, data, deriv[None, :], 1)
