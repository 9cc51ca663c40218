"""Objective-function layer over the GPU calculators, with batched evaluation.

* ThreeDModelObjective mirrors jif3D's ObjectiveFunction::CalcMisfit / CalcGradient
  (jif3D/Inversion/ObjectiveFunction.h:123-229) with ThreeDModelObjective's data
  difference and gradient hook (jif3D/Inversion/ThreeDModelObjective.h:211-329): chi^2 =
  sum_i d_i^2 / e_i^2 with d = synthetic - observed (zero where |observed| <=
  ethresh * error), and the gradient LQDerivative(model, d / e^2).
* CalcMisfitBatch / CalcGradientBatch evaluate several model vectors at once (e.g. the
  trial steps of a line search, which jif3D's mcsrch evaluates one at a time,
  jif3D/Inversion/mcsrch.cpp:247-248): the calculator's batched forward and gradient
  share the corner-function evaluations (matrix-free FP64) or the pass over the stored J
  (FullSensitivityGravMagCalculator).
* TanhTransform mirrors jif3D/ModelTransforms/TanhTransform.h:71-114 (generalized <->
  physical parameters and the gradient chain rule that ModelDistributor::TransformGradient
  applies, jif3D/Inversion/ModelDistributor.h:89-94).
"""
import numpy as np


class ThreeDModelObjective:
    def __init__(self, calculator, ethresh=0.0):
        self.Calculator = calculator
        self.ignorelargeerror = float(ethresh)
        self.ObservedData = None
        self.CoarseModel = None
        self.InvCovDiag = None
        self.DataDifference = None
        self.IndividualMisfits = None
        self.nEval = 0
        self._last_model = None
        self._batch_models = None
        self._batch_diff = None

    def SetDataTransform(self, transform):
        self.Calculator.SetDataTransform(transform)

    def SetObservedData(self, data):
        if np.asarray(data.GetData()).size == 0:
            raise ValueError("Cannot have empty observations in objective function. ")
        self.ObservedData = data
        self.SetDataError(data.GetErrors())

    def SetDataError(self, errors):
        e = np.asarray(errors, dtype=np.float64)
        self.InvCovDiag = 1.0 / (e * e)

    def GetDataError(self):
        return 1.0 / np.sqrt(self.InvCovDiag)

    def SetCoarseModelGeometry(self, model):
        if model.ncells == 0:
            raise ValueError("Cannot have empty model in objective function. ")
        self.CoarseModel = model

    def _check_model(self, m):
        m = np.asarray(m, dtype=np.float64)
        if m.shape[-1] < self.CoarseModel.ncells:
            raise ValueError("Not enough inversion parameters to fill the grid !")
        return m

    def _difference(self, synth):
        obs = np.asarray(self.ObservedData.GetData(), dtype=np.float64)
        if synth.shape[-1] != obs.size:
            raise ValueError(" ThreeDModelObjective: Forward calculation does not give same "
                             "amount of data !")
        err = self.GetDataError()
        keep = np.abs(obs) > self.ignorelargeerror * err
        return np.where(keep, synth - obs, 0.0)

    def _weights(self, ndata):
        if self.InvCovDiag is None or self.InvCovDiag.size != ndata:
            self.InvCovDiag = np.ones(ndata)
        return self.InvCovDiag

    def CalcMisfit(self, m):
        """chi^2 misfit of one model vector; stores the weighted difference for
        CalcGradient (ObjectiveFunction.h:165-200)."""
        m = self._check_model(m)
        self.CoarseModel.SetModelParameters(m)
        synth = self.Calculator.Calculate(self.CoarseModel, self.ObservedData)
        diff = self._difference(synth)
        tmp = self._weights(diff.size) * diff
        self.nEval += 1
        self.IndividualMisfits = np.sqrt(diff * tmp) * np.sign(diff)
        self.DataDifference = tmp
        self._last_model = m.copy()
        return float(np.sum(diff * tmp))

    def CalcGradient(self, m):
        """Gradient of the last CalcMisfit's model (ThreeDModelObjective.h:288-329)."""
        m = self._check_model(m)
        if self._last_model is None or not np.array_equal(m[: self.CoarseModel.ncells],
                                                          self._last_model[: self.CoarseModel.ncells]):
            raise ValueError("Gradient calculation needs identical model to forward !")
        self.nEval += 1
        return self.Calculator.LQDerivative(self.CoarseModel, self.ObservedData,
                                            self.DataDifference)

    def CalcMisfitBatch(self, models):
        """chi^2 of every row of `models` in one batched forward; stores the weighted
        differences for CalcGradientBatch."""
        models = np.atleast_2d(self._check_model(models))
        synth = self.Calculator.CalculateBatch(self.CoarseModel, self.ObservedData, models)
        diff = self._difference(synth)
        tmp = self._weights(diff.shape[1])[None, :] * diff
        self.nEval += models.shape[0]
        self._batch_models = models.copy()
        self._batch_diff = tmp
        return np.sum(diff * tmp, axis=1)

    def CalcGradientBatch(self, models):
        models = np.atleast_2d(self._check_model(models))
        if self._batch_models is None or not np.array_equal(models, self._batch_models):
            raise ValueError("Gradient calculation needs identical models to forward !")
        self.nEval += models.shape[0]
        return self.Calculator.LQDerivativeBatch(self.CoarseModel, self.ObservedData, models,
                                                 self._batch_diff)

    def EvaluateAlongDirection(self, m, direction, steps):
        """Misfits and gradients of the trial models m + step * direction for every
        step, in one batch each: the evaluations a line search needs for several step
        lengths."""
        m = np.asarray(m, dtype=np.float64)
        p = np.asarray(direction, dtype=np.float64)
        trials = m[None, :] + np.asarray(steps, dtype=np.float64)[:, None] * p[None, :]
        return self.CalcMisfitBatch(trials), self.CalcGradientBatch(trials)


class TanhTransform:
    """Physical parameters m = min + (1 + tanh x) / 2 (max - min) of generalized ones x."""

    def __init__(self, minval, maxval):
        self.min = np.atleast_1d(np.asarray(minval, dtype=np.float64))
        self.max = np.atleast_1d(np.asarray(maxval, dtype=np.float64))

    def _bounds(self, n):
        if self.min.size == n and self.max.size == n:
            return self.min, self.max
        if self.min.size == 1 and self.max.size == 1:
            return np.full(n, self.min[0]), np.full(n, self.max[0])
        raise ValueError("TanhTransform: bounds and model differ in size")

    def GeneralizedToPhysical(self, x):
        x = np.asarray(x, dtype=np.float64)
        lo, hi = self._bounds(x.shape[-1])
        return lo + (1.0 + np.tanh(x)) / 2.0 * (hi - lo)

    def PhysicalToGeneralized(self, m):
        m = np.asarray(m, dtype=np.float64)
        lo, hi = self._bounds(m.shape[-1])
        return np.arctanh(2.0 * (m - lo) / (hi - lo) - 1.0)

    def Derivative(self, x, gradient):
        """Gradient with respect to x from the gradient with respect to m."""
        x = np.asarray(x, dtype=np.float64)
        lo, hi = self._bounds(x.shape[-1])
        return (hi - lo) / (2.0 * np.cosh(x) ** 2) * np.asarray(gradient, dtype=np.float64)
