"""Implementation and calculator classes mirroring jif3D's GravMag bridge.

* ScalarHipGravityImp, TensorHipGravityImp, HipMagneticSusceptibilityImp and
  ScalarTensorHipGravityImp play the role of ScalarOMPGravityImp /
  TensorOMPGravityImp / OMPMagneticSusceptibilityImp
  (jif3D/Gravity/ScalarOMPGravityImp.cpp:124-282, TensorOMPGravityImp.cpp:112-267,
  jif3D/Magnetics/OMPMagneticSusceptibilityImp.cpp:31-201): Calculate() and
  LQDerivative() of ThreeDGravMagImplementation (jif3D/GravMag/
  ThreeDGravMagImplementation.h:116-231) run on the GPU through the C ABI.
* MinMemGravMagCalculator mirrors jif3D/GravMag/MinMemGravMagCalculator.h:25-79;
  FullSensitivityGravMagCalculator mirrors jif3D/GravMag/FullSensitivityGravMagCalculator.h
  with the geometry cache of CachedGravMagCalculator.h:27-135, keeping J in HBM.
* FTGInvariant / TotalField / TotalFieldAnomaly select the per-site transforms of
  jif3D/Gravity/GravityTransforms.h:25-100 and jif3D/Magnetics/MagneticTransforms.h:19-156.

Arrays are numpy on the host; the device work happens inside libjif3d_amd.so.
"""
import numpy as np

from . import _native as nat


class FTGInvariant:
    code = nat.TF_FTG_INVARIANT
    angles = (0.0, 0.0)

    def GetInputSize(self):
        return 9

    def GetOutputSize(self):
        return 1


class TotalField:
    code = nat.TF_TOTAL_FIELD
    angles = (0.0, 0.0)

    def GetInputSize(self):
        return 3

    def GetOutputSize(self):
        return 1


class TotalFieldAnomaly:
    code = nat.TF_TOTAL_FIELD_ANOMALY

    def __init__(self, inc=0.0, dec=0.0):
        self.angles = (float(inc), float(dec))

    def GetInputSize(self):
        return 3

    def GetOutputSize(self):
        return 1


def _geometry_key(model, data):
    parts = list(model.coords) + list(model.sizes) + [np.asarray(model.bg_thicknesses)]
    parts += list(data.positions)
    return tuple(np.asarray(p, dtype=np.float64).tobytes() for p in parts)


class HipGravMagImplementation:
    """GPU implementation object; kind and precision fixed at construction."""

    kind = None

    def __init__(self, precision=nat.FP64, device=0, stream=None):
        self.precision = precision
        self.Transform = None
        self._ctx = None
        self._key = None
        self._device = device
        self._stream = stream

    def RawDataPerMeasurement(self):
        return nat.RAW_PER_SITE[self.kind]

    def SetDataTransform(self, transform):
        self.Transform = transform

    def GetDataPerMeasurement(self):
        return self.Transform.GetOutputSize() if self.Transform else self.RawDataPerMeasurement()

    def _field(self):
        return None

    def context(self, model, data):
        if self._ctx is None:
            self._ctx = nat.Context(self._device, self._stream)
        key = _geometry_key(model, data)
        if key != self._key:
            self._ctx.configure(model, data, self._field())
            self._key = key
        elif self._field() is not None:
            self._ctx.set_field(*self._field())
        tf = self.Transform
        self._ctx.set_transform(tf.code if tf else nat.TF_NONE, *(tf.angles if tf else (0, 0)))
        return self._ctx

    def Calculate(self, model, data):
        ctx = self.context(model, data)
        return ctx.forward(self.kind, self.precision, model.GetModelParameters())

    def LQDerivative(self, model, data, misfit):
        ctx = self.context(model, data)
        misfit = np.asarray(misfit, dtype=np.float64)
        need = data.nsites * self.GetDataPerMeasurement()
        if misfit.size != need:
            raise ValueError(f"misfit has {misfit.size} entries, expected {need}")
        return ctx.adjoint(self.kind, self.precision, model.GetModelParameters(), misfit)

    def CalculateBatch(self, model, data, models):
        """Calculate for several model vectors on the geometry of `model` (rows of
        `models`); FP64 batches share the corner-function evaluations."""
        ctx = self.context(model, data)
        return ctx.forward_batch(self.kind, self.precision, models)

    def LQDerivativeBatch(self, model, data, models, misfits):
        """LQDerivative for several (model vector, misfit) pairs, one per row."""
        ctx = self.context(model, data)
        misfits = np.asarray(misfits, dtype=np.float64)
        need = data.nsites * self.GetDataPerMeasurement()
        if misfits.ndim != 2 or misfits.shape[1] != need:
            raise ValueError(f"misfits must have shape (B, {need})")
        return ctx.adjoint_batch(self.kind, self.precision, models, misfits)

    def SensitivityRows(self, model, data, site0=0, nsites=None):
        """Raw FP64 J rows (reference layout) of a site range, copied to the host."""
        ctx = self.context(model, data)
        nsites = data.nsites - site0 if nsites is None else nsites
        k = self.RawDataPerMeasurement()
        out = ctx.alloc((nsites * k, ctx.model_size(self.kind)))
        ctx.sensitivity_rows(self.kind, site0, nsites, out)
        ctx.synchronize()
        host = out.to_host()
        out.free()
        return host


class ScalarHipGravityImp(HipGravMagImplementation):
    kind = nat.KIND_SCALAR


class TensorHipGravityImp(HipGravMagImplementation):
    kind = nat.KIND_TENSOR


class ScalarTensorHipGravityImp(HipGravMagImplementation):
    """gz and the 9 FTG values per site in one pass (10 raw values per site)."""
    kind = nat.KIND_SCALAR_TENSOR


class HipMagneticSusceptibilityImp(HipGravMagImplementation):
    """OMPMagneticSusceptibilityImp(Inc, Dec, Fs): radians, radians, nT."""
    kind = nat.KIND_MAGNETIC

    def __init__(self, Inc, Dec, Fs, precision=nat.FP64, device=0, stream=None):
        super().__init__(precision, device, stream)
        self.Inclination, self.Declination, self.FieldStrength = float(Inc), float(Dec), float(Fs)

    def _field(self):
        return (self.Inclination, self.Declination, self.FieldStrength)


class HipMagneticGradImp(HipGravMagImplementation):
    """OMPMagneticGradImp(Inc, Dec, Fs) (jif3D/Magnetics/OMPMagneticGradImp.cpp:13-118):
    one value per site, Bz at the site minus Bz 1 m below, no background.  FP64 by
    default; nat.MIXED evaluates the difference as the field of 1 m slabs at the mesh
    planes (FP64 when a cell is thinner than 1 m)."""
    kind = nat.KIND_MAGNETIC_GRADIENT

    def __init__(self, Inc=0.0, Dec=0.0, Fs=1.0, precision=nat.FP64, device=0, stream=None):
        super().__init__(precision, device, stream)
        self.Inclination, self.Declination, self.FieldStrength = float(Inc), float(Dec), float(Fs)

    def _field(self):
        return (self.Inclination, self.Declination, self.FieldStrength)


class HipMagnetizationImp(HipGravMagImplementation):
    """OMPMagnetizationImp (jif3D/Magnetics/OMPMagnetizationImp.cpp:25-125): Bx, By, Bz
    of a ThreeDMagnetizationModel (model vector [Mx; My; Mz], no background).  FP64 by
    default; nat.MIXED runs the tensor far field once per pair for the three components."""
    kind = nat.KIND_MAGNETIZATION

    def __init__(self, precision=nat.FP64, device=0, stream=None):
        super().__init__(precision, device, stream)


class ThreeDGravMagCalculator:
    """Calculator base (jif3D/GravMag/ThreeDGravMagCalculator.h:48-173)."""

    def __init__(self, imp):
        self.Imp = imp
        self.Transform = None

    def SetDataTransform(self, transform):
        self.Transform = transform
        self.Imp.SetDataTransform(transform)

    def GetDataPerMeasurement(self):
        return self.Imp.GetDataPerMeasurement()

    def CheckModelConsistency(self, model):
        model.CheckConsistency()

    def Calculate(self, model, data):
        self.CheckModelConsistency(model)
        return self.Imp.Calculate(model, data)

    def LQDerivative(self, model, data, misfit):
        self.CheckModelConsistency(model)
        return self.Imp.LQDerivative(model, data, misfit)

    def CalculateBatch(self, model, data, models):
        """Forward responses of several model vectors (rows) on the geometry of `model`:
        the batched form of Calculate for line-search trials."""
        self.CheckModelConsistency(model)
        return self.Imp.CalculateBatch(model, data, models)

    def LQDerivativeBatch(self, model, data, models, misfits):
        self.CheckModelConsistency(model)
        return self.Imp.LQDerivativeBatch(model, data, models, misfits)


class MinMemGravMagCalculator(ThreeDGravMagCalculator):
    """No stored sensitivities: forward and gradient are recomputed matrix-free."""
    pass


class CachedGravMagCalculator(ThreeDGravMagCalculator):
    """Geometry cache (jif3D/GravMag/CachedGravMagCalculator.h:27-135): while mesh,
    background layers, sites (and inducing field) are unchanged, Calculate and
    LQDerivative use the cached quantities of the derived class."""

    def __init__(self, imp):
        super().__init__(imp)
        self._cache_key = None

    def _key(self, model, data):
        return _geometry_key(model, data) + (repr(self.Imp._field()),)

    def InvalidateCache(self):
        self._cache_key = None

    def Calculate(self, model, data):
        self.CheckModelConsistency(model)
        key = self._key(model, data)
        if self._cache_key == key:
            return self.CalculateCachedResult(model, data)
        result = self.CalculateNewModel(model, data)
        self._cache_key = key
        return result

    def LQDerivative(self, model, data, misfit):
        self.CheckModelConsistency(model)
        if self._cache_key == self._key(model, data):
            return self.CachedLQDerivative(model, data, misfit)
        return self.Imp.LQDerivative(model, data, misfit)

    def CalculateBatch(self, model, data, models):
        """With a valid cache every model goes through the cached path in one batch;
        otherwise the first model is calculated as new (filling the cache, as
        Calculate would) and the others through the cache."""
        self.CheckModelConsistency(model)
        models = np.atleast_2d(np.asarray(models, dtype=np.float64))
        key = self._key(model, data)
        if self._cache_key == key:
            return self.CalculateCachedResultBatch(model, data, models)
        saved = model.GetModelParameters().copy()
        model.SetModelParameters(models[0])
        try:
            first = self.CalculateNewModel(model, data)
        finally:
            model.SetModelParameters(saved)
        self._cache_key = key
        rest = (self.CalculateCachedResultBatch(model, data, models[1:]) if len(models) > 1
                else np.zeros((0, first.size)))
        return np.vstack([first[None, :], rest])

    def LQDerivativeBatch(self, model, data, models, misfits):
        self.CheckModelConsistency(model)
        if self._cache_key == self._key(model, data):
            return self.CachedLQDerivativeBatch(model, data, models, misfits)
        return self.Imp.LQDerivativeBatch(model, data, models, misfits)


class FullSensitivityGravMagCalculator(CachedGravMagCalculator):
    """Stored sensitivities (jif3D/GravMag/FullSensitivityGravMagCalculator.h:99-221),
    kept in HBM in the compact row layout: the first Calculate for a geometry builds J
    and returns the imp's direct result; later calls are J m (plus the background layers)
    and 2 [J B]^T D^T misfit on the matrix cores in `gemm_precision`.  J is FP32, or FP64
    (the reference's precision) with gemm_precision=GEMM_FP64."""

    host_block_cells = None

    def __init__(self, imp, gemm_precision=nat.GEMM_FP32):
        super().__init__(imp)
        self.gemm_precision = gemm_precision

    def CalculateNewModel(self, model, data):
        ctx = self.Imp.context(model, data)
        ctx.storedj_build(self.Imp.kind, self.Imp.precision, self.host_block_cells,
                          fp64=self.gemm_precision == nat.GEMM_FP64)
        return self.Imp.Calculate(model, data)

    def CalculateCachedResult(self, model, data):
        ctx = self.Imp.context(model, data)
        return ctx.storedj_forward(self.gemm_precision, model.GetModelParameters())

    def CachedLQDerivative(self, model, data, misfit):
        ctx = self.Imp.context(model, data)
        misfit = np.asarray(misfit, dtype=np.float64)
        need = data.nsites * self.GetDataPerMeasurement()
        if misfit.size != need:
            raise ValueError(f"misfit has {misfit.size} entries, expected {need}")
        return ctx.storedj_gradient(self.gemm_precision, model.GetModelParameters(), misfit)

    def CalculateCachedResultBatch(self, model, data, models):
        ctx = self.Imp.context(model, data)
        return ctx.storedj_forward_batch(self.gemm_precision, models)

    def CachedLQDerivativeBatch(self, model, data, models, misfits):
        ctx = self.Imp.context(model, data)
        misfits = np.asarray(misfits, dtype=np.float64)
        need = data.nsites * self.GetDataPerMeasurement()
        if misfits.ndim != 2 or misfits.shape[1] != need:
            raise ValueError(f"misfits must have shape (B, {need})")
        return ctx.storedj_gradient_batch(self.gemm_precision, models, misfits)


class DiskGravMagCalculator(FullSensitivityGravMagCalculator):
    """The role of jif3D/GravMag/DiskGravMagCalculator.h:100-200 (a J larger than the
    memory that computes with it, kept elsewhere and read back for every cached product):
    J is kept in page-locked host memory, or with `TempDir` in a file in that directory as
    the reference keeps it (DiskGravMagCalculator(TheImp, TDir), :64-66), and streamed to
    HBM in blocks of `block_cells` cells, each block's copy (and file read) overlapping the
    products on the previous one.  The cached gradient keeps the factor 2 that the
    reference's Disk calculator omits."""

    def __init__(self, imp, gemm_precision=nat.GEMM_FP32, block_cells=16384, TempDir=None):
        super().__init__(imp, gemm_precision)
        self.host_block_cells = int(block_cells)
        self.TempDir = TempDir

    def CalculateNewModel(self, model, data):
        if self.TempDir is None:
            return super().CalculateNewModel(model, data)
        ctx = self.Imp.context(model, data)
        ctx.storedj_build(self.Imp.kind, self.Imp.precision, self.host_block_cells,
                          directory=self.TempDir)
        return self.Imp.Calculate(model, data)
