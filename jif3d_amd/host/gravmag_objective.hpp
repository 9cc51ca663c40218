// Objective function over the adapter's calculators, with batched evaluation.
//
// GravMagObjective<Calc> mirrors ObjectiveFunction::CalcMisfit / CalcGradient
// (jif3D/Inversion/ObjectiveFunction.h:123-229) with ThreeDModelObjective's data
// difference and gradient hook (jif3D/Inversion/ThreeDModelObjective.h:211-329):
// chi^2 = sum_i d_i^2 / e_i^2, d = synthetic - observed, zero where |observed| <=
// ethresh * error; the gradient is LQDerivative(model, d / e^2) at the model of the
// last CalcMisfit.  CalcMisfitBatch / CalcGradientBatch do the same for several model
// vectors in one batched forward and one batched gradient (the trial steps of a line
// search, which jif3D/Inversion/mcsrch.cpp:247-248 evaluates one at a time).
//
// Calc is HipGravMagImplementation or HipFullSensitivityCalculator
// (hip_gravmag_imp.hpp); the mesh and sites are fixed at construction.
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

#include "hip_gravmag_imp.hpp"

namespace jif3d_amd {

template <class Calc>
class GravMagObjective {
 public:
  GravMagObjective(Calc& calc, Mesh mesh, Sites sites, double ethresh = 0.0)
      : calc_(calc), mesh_(std::move(mesh)), sites_(std::move(sites)), ethresh_(ethresh) {}

  // ThreeDModelObjective::SetObservedData with GeneralData::SetDataAndErrors.
  void SetObservedData(const std::vector<double>& data, const std::vector<double>& errors) {
    if (data.empty()) throw FatalException("Cannot have empty observations in objective function.");
    if (data.size() != errors.size())
      throw FatalException("Data and Errors do not have the same size !");
    observed_ = data;
    errors_ = errors;
    invcov_.resize(errors.size());
    for (size_t i = 0; i < errors.size(); ++i) invcov_[i] = 1.0 / (errors[i] * errors[i]);
  }

  double CalcMisfit(const std::vector<double>& model) {
    const std::vector<double> synth = calc_.Calculate(mesh_, sites_, model);
    last_model_ = model;
    return Weigh(synth.data(), 1, &diff_)[0];
  }

  std::vector<double> CalcGradient(const std::vector<double>& model) {
    if (model != last_model_)
      throw FatalException("Gradient calculation needs identical model to forward !");
    return calc_.LQDerivative(mesh_, sites_, model, diff_);
  }

  // models: nmodels model vectors, model-major; returns one chi^2 per model.
  std::vector<double> CalcMisfitBatch(const std::vector<double>& models, size_t nmodels) {
    const std::vector<double> synth = calc_.CalculateBatch(mesh_, sites_, models, nmodels);
    batch_models_ = models;
    return Weigh(synth.data(), nmodels, &batch_diff_);
  }

  // Gradients of the last CalcMisfitBatch's models: nmodels * model length values.
  std::vector<double> CalcGradientBatch(const std::vector<double>& models, size_t nmodels) {
    if (models != batch_models_)
      throw FatalException("Gradient calculation needs identical models to forward !");
    return calc_.LQDerivativeBatch(mesh_, sites_, models, batch_diff_, nmodels);
  }

 private:
  // chi^2 of each of nb synthetic data vectors; stores the weighted differences.
  std::vector<double> Weigh(const double* synth, size_t nb, std::vector<double>* wdiff) const {
    const size_t nd = observed_.size();
    if (nd == 0) throw FatalException("no observed data (SetObservedData)");
    std::vector<double> chi(nb, 0.0);
    wdiff->assign(nb * nd, 0.0);
    for (size_t b = 0; b < nb; ++b)
      for (size_t i = 0; i < nd; ++i) {
        const double d = std::abs(observed_[i]) > ethresh_ * errors_[i]
                             ? synth[b * nd + i] - observed_[i]
                             : 0.0;
        const double w = invcov_[i] * d;
        (*wdiff)[b * nd + i] = w;
        chi[b] += d * w;
      }
    return chi;
  }

  Calc& calc_;
  Mesh mesh_;
  Sites sites_;
  double ethresh_;
  std::vector<double> observed_, errors_, invcov_;
  std::vector<double> last_model_, diff_, batch_models_, batch_diff_;
};

}  // namespace jif3d_amd
