// Depth weighting of the sensitivities (jif3D/Gravity/DepthWeighting.h:24-133,
// DepthWeighting.cpp:14-160) on std::vector: WeightingTerm f(z, z0) = (z + z0)^n, the
// one-parameter fit of z0 to a sensitivity profile, the weight vector 1 / f normalised
// to 1, and the profile below a site at the middle of the mesh from the adapter's
// matrix-free adjoint (one site, unit residual).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

#include "hip_gravmag_imp.hpp"

namespace jif3d_amd {

class WeightingTerm {
 public:
  explicit WeightingTerm(double exponent) : n_(exponent) {}
  double operator()(double z, double z0) const { return std::pow(z + z0, n_); }
  double deriv(double z, double z0) const { return n_ * std::pow(z + z0, n_ - 1.0); }
  double average(double z1, double z2, double z0) const {
    return 1.0 / ((z2 - z1) * (n_ + 1.0)) *
           (std::pow(z2 + z0, n_ + 1.0) - std::pow(z1 + z0, n_ + 1.0));
  }

 private:
  double n_;
};

// Gauss-Newton fit of z0 (DepthWeighting.cpp:52-116): start at 1.1 times the mesh
// depth, at most 100 steps, stop when a step is below 0.01 m.  The reference's
// one-parameter ModelSpaceInversion with unit errors and lambda = 0 is the step
// (s^T s)^-1 s^T delta.  No "match.out" file is written.
inline double FitZ0(const std::vector<double>& sens_profile, const std::vector<double>& zsizes,
                    const WeightingTerm& f) {
  const size_t nz = sens_profile.size();
  if (nz == 0 || zsizes.size() != nz) throw FatalException("FitZ0: profile and sizes differ");
  std::vector<double> profile(nz), zvalues(nz), calc(nz), sens(nz);
  double acc = 0.0;
  for (size_t k = 0; k < nz; ++k) {
    acc += zsizes[k];
    zvalues[k] = acc;
    profile[k] = std::abs(sens_profile[k]);
  }
  const double maxel = *std::max_element(profile.begin(), profile.end());
  for (double& v : profile) v /= maxel;
  double z0 = zvalues[nz - 1] * 1.1;
  double stepsize = 1e6;
  for (size_t it = 0; stepsize > 0.01 && it < 100; ++it) {
    for (size_t k = 0; k < nz; ++k) calc[k] = f(zvalues[k], z0);
    const double maximum = *std::max_element(calc.begin(), calc.end());
    double sd = 0.0, ss = 0.0;
    for (size_t k = 0; k < nz; ++k) {
      calc[k] /= maximum;
      sens[k] = f.deriv(zvalues[k], z0) / maximum;
      sd += sens[k] * (profile[k] - calc[k]);
      ss += sens[k] * sens[k];
    }
    const double step = sd / ss;
    stepsize = std::abs(step);
    z0 -= step;
  }
  return z0;
}

// 1 / f(bottom of layer k, z0), normalised to a maximum of 1 (DepthWeighting.cpp:128-152).
inline std::vector<double> ConstructDepthWeighting(const std::vector<double>& zsizes, double z0,
                                                   const WeightingTerm& f) {
  std::vector<double> w(zsizes.size());
  double bottom = 0.0;
  for (size_t k = 0; k < zsizes.size(); ++k) {
    bottom += zsizes[k];
    w[k] = 1.0 / f(bottom, z0);
  }
  const double maxel = *std::max_element(w.begin(), w.end());
  for (double& v : w) v /= maxel;
  return w;
}

// Index of the cell of `coords` (n+1 corner coordinates) containing v, clamped at 0
// (ThreeDModelBase::FindAssociatedIndices, ThreeDModelBase.cpp:137-156).
inline size_t AssociatedIndex(const std::vector<double>& coords, double v) {
  const auto it = std::lower_bound(coords.begin(), coords.end(), v);
  const long idx = static_cast<long>(it - coords.begin()) - 1;
  return idx > 0 ? static_cast<size_t>(idx) : 0;
}

// The nz gradient values below a site at (last x coordinate / 2, last y coordinate / 2,
// -0.01 m) for a unit residual of every datum (CalculateMiddleSens, DepthWeighting.h:109-133).
template <class Calc>
std::vector<double> CalculateMiddleSens(Calc& calc, const Mesh& mesh,
                                        const std::vector<double>& model) {
  Sites site;
  site.x = {mesh.xcoords.back() / 2.0};
  site.y = {mesh.ycoords.back() / 2.0};
  site.z = {-0.01};
  calc.Calculate(mesh, site, model);
  const std::vector<double> misfit(calc.GetDataPerMeasurement(), 1.0);
  const std::vector<double> deriv = calc.LQDerivative(mesh, site, model, misfit);
  const size_t ny = mesh.ysizes.size(), nz = mesh.zsizes.size();
  const size_t i = AssociatedIndex(mesh.xcoords, site.x[0]);
  const size_t j = AssociatedIndex(mesh.ycoords, site.y[0]);
  const size_t start = (nz * ny) * i + nz * j;
  return std::vector<double>(deriv.begin() + start, deriv.begin() + start + nz);
}

}  // namespace jif3d_amd
