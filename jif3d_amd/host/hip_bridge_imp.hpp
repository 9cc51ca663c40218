// HipGravMagImp: the jif3D imp class over libjif3d_amd.so, written against jif3D's own
// bridge (ThreeDGravMagImplementation, jif3D/GravMag/ThreeDGravMagImplementation.h:39-231)
// so that it drops in behind the unchanged calculators.  Include it after the jif3D
// headers that define ThreeDGravMagImplementation, ThreeDGravMagCalculator, rvec, rmat and
// FatalException (in the jif3D tree), or after jif3d_bridge.hpp (the same classes on
// std::vector, for building and testing outside it).
//
// Two contracts of the base class shape it:
// * The data transform is private to the base (:46) and SetDataTransform is not virtual
//   (:67-70), so the imp never sees which transform the calculator set.  The base class's
//   own site loop applies it: Calculate hands the loop raw data (and sensitivity rows)
//   site by site from one GPU batch, and LQDerivative obtains each site's D^T r (D the
//   transform's Jacobian at the site's raw data) from the base class's LQDerivative run on
//   a one-site probe whose sensitivity rows are the identity, then calls the matrix-free
//   GPU adjoint with those raw residuals.
// * Calculators that store J (FullSensitivity, Disk) size SetCurrentSensitivities() to
//   RawDataPerMeasurement() x GetNModelParm() and copy it in HandleSensitivities(i) after
//   each site (CachedGravMagCalculator.h:113-115, FullSensitivityGravMagCalculator.h:
//   99-114); MinMem sizes it 0 x 0 (MinMemGravMagCalculator.h:76).  When it is sized (the
//   storesens test of ScalarOMPGravityImp.cpp:150-151), CalcGridded fills it with the FP64
//   rows of the site, gridded and background columns, fetched from the GPU in blocks of
//   sites (jgm_sensitivity_rows); CalcBackground then returns zero, its part being in the
//   rows and the raw data already.
// Without a transform and without stored rows, Calculate is one whole-survey GPU call.
#pragma once

#include <algorithm>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include "jif3d_gravmag.h"

namespace jif3D {

template <class DataType>
class HipGravMagImp : public ThreeDGravMagImplementation<DataType> {
 public:
  typedef ThreeDGravMagImplementation<DataType> Base;
  typedef typename Base::ThreeDModelType ModelType;

  HipGravMagImp(jgm_kind kind, jgm_precision prec = JGM_FP64, int device = 0)
      : kind_(kind), prec_(prec) {
    if (jgm_api_version() != JGM_API_VERSION)
      throw FatalException("libjif3d_amd.so was built from another C ABI version", __FILE__,
                           __LINE__);
    Check(jgm_create(device, nullptr, &ctx_));
    const size_t k = jgm_raw_per_site(kind);
    probe_model_.SetMeshSize(1, 1, k);
    probe_data_.AddMeasurementPoint(0.0, 0.0, 0.0);
  }
  ~HipGravMagImp() override {
    if (rows_dev_) jgm_device_free(ctx_, rows_dev_);
    jgm_destroy(ctx_);
  }
  HipGravMagImp(const HipGravMagImp&) = delete;
  HipGravMagImp& operator=(const HipGravMagImp&) = delete;

  // inducing field of the susceptibility and vertical-gradient kinds (radians, radians, nT),
  // the constructor arguments of OMPMagneticSusceptibilityImp
  void SetField(double inclination, double declination, double fieldstrength) {
    Check(jgm_set_field(ctx_, inclination, declination, fieldstrength));
  }
  size_t RawDataPerMeasurement() override { return jgm_raw_per_site(kind_); }

  rvec Calculate(const ModelType& Model, const DataType& Data,
                 ThreeDGravMagCalculator<DataType>& Calculator) override {
    const size_t S = Configure(Model, Data);
    const size_t k = RawDataPerMeasurement();
    const rvec m = Model.GetModelParameters();
    RawData(m, S);
    rmat& cur = Calculator.SetCurrentSensitivities();
    const bool store_rows = cur.size1() >= k && cur.size2() >= Model.GetNModelParm();
    if (!store_rows && !HasTransform()) {
      rvec out(S * k);
      std::copy(raw_.begin(), raw_.end(), out.begin());
      return out;
    }
    // the base class's site loop applies the transform and hands the rows over
    Serving guard(this, Model.GetNModelParm());
    return Base::Calculate(Model, Data, Calculator);
  }

  rvec LQDerivative(const ModelType& Model, const DataType& Data, const rvec& Misfit) override {
    const size_t S = Configure(Model, Data);
    const size_t k = RawDataPerMeasurement(), kt = this->GetDataPerMeasurement();
    if (Misfit.size() != S * kt)
      throw FatalException("misfit length does not match the data", __FILE__, __LINE__);
    const rvec m = Model.GetModelParameters();
    std::vector<double> rm(S * k);
    if (HasTransform()) {
      RawData(m, S);
      for (size_t i = 0; i < S; ++i) {
        const std::vector<double> a = TransformAdjoint(&raw_[i * k], Misfit, i * kt);
        std::copy(a.begin(), a.end(), rm.begin() + i * k);
      }
    } else {
      for (size_t i = 0; i < S * k; ++i) rm[i] = Misfit(i);
    }
    rvec grad(Model.GetNModelParm());
    if (S > 0)
      Check(jgm_adjoint_host(ctx_, kind_, prec_, Ptr(m), rm.data(), &grad(0)));
    else
      for (size_t c = 0; c < grad.size(); ++c) grad(c) = 0.0;
    return grad;
  }

 private:
  // ---- the base class's per-site hooks --------------------------------------------------
  rvec CalcGridded(const size_t measindex, const ModelType&, const DataType&,
                   rmat& Sensitivities) override {
    const size_t k = RawDataPerMeasurement();
    rvec out(k);
    if (probe_) {
      for (size_t j = 0; j < k; ++j) out(j) = probe_raw_[j];
      for (size_t r = 0; r < Sensitivities.size1(); ++r)
        for (size_t c = 0; c < Sensitivities.size2(); ++c) Sensitivities(r, c) = r == c;
      return out;
    }
    if (!serving_)
      throw FatalException("HipGravMagImp evaluates sites only inside its Calculate", __FILE__,
                           __LINE__);
    for (size_t j = 0; j < k; ++j) out(j) = raw_[measindex * k + j];
    if (Sensitivities.size1() >= k && Sensitivities.size2() >= nmod_)
      FillRows(measindex, Sensitivities);
    return out;
  }
  rvec CalcBackground(const size_t, const double, const double, const double, const ModelType&,
                      const DataType&, rmat&) override {
    rvec z(RawDataPerMeasurement());
    for (size_t j = 0; j < z.size(); ++j) z(j) = 0.0;
    return z;
  }

  // ---- helpers ------------------------------------------------------------------------
  static void Check(jgm_status st) {
    if (st != JGM_OK) throw FatalException(jgm_last_error(), __FILE__, __LINE__);
  }
  static const double* Ptr(const rvec& v) { return v.size() ? &v(0) : nullptr; }
  template <class V>
  static const double* Ptr(const V& v) { return v.empty() ? nullptr : &v[0]; }

  // background layers of the models that have them (the magnetization model has none)
  template <class M>
  static auto Background(const M& model, int)
      -> decltype(model.GetBackgroundThicknesses(), std::vector<double>()) {
    const auto& t = model.GetBackgroundThicknesses();
    return std::vector<double>(t.begin(), t.end());
  }
  template <class M>
  static std::vector<double> Background(const M&, long) {
    return std::vector<double>();
  }

  // mesh, background and sites to the library when they changed; returns nsites
  size_t Configure(const ModelType& Model, const DataType& Data) {
    std::vector<double> key;
    const std::vector<double> bg = Background(Model, 0);
    const std::vector<double>* axes[] = {&Model.GetXCoordinates(), &Model.GetYCoordinates(),
                                         &Model.GetZCoordinates(), &Model.GetXCellSizes(),
                                         &Model.GetYCellSizes(), &Model.GetZCellSizes(), &bg};
    for (const auto* v : axes) {
      key.push_back(static_cast<double>(v->size()));
      key.insert(key.end(), v->begin(), v->end());
    }
    for (const auto* v : {&Data.GetMeasPosX(), &Data.GetMeasPosY(), &Data.GetMeasPosZ()})
      key.insert(key.end(), v->begin(), v->end());
    const size_t S = Data.GetMeasPosX().size();
    if (key != geometry_) {
      Check(jgm_set_mesh(ctx_, Model.GetXCellSizes().size(), Model.GetYCellSizes().size(),
                         Model.GetZCellSizes().size(), Ptr(Model.GetXCoordinates()),
                         Ptr(Model.GetYCoordinates()), Ptr(Model.GetZCoordinates()),
                         Ptr(Model.GetXCellSizes()), Ptr(Model.GetYCellSizes()),
                         Ptr(Model.GetZCellSizes())));
      Check(jgm_set_background(ctx_, bg.size(), Ptr(bg)));
      Check(jgm_set_sites(ctx_, S, Ptr(Data.GetMeasPosX()), Ptr(Data.GetMeasPosY()),
                          Ptr(Data.GetMeasPosZ())));
      geometry_.swap(key);
    }
    if (jgm_model_size(ctx_, kind_) != Model.GetNModelParm())
      throw FatalException("model vector length does not match the mesh", __FILE__, __LINE__);
    // the library computes raw data; the base class applies the transform
    Check(jgm_set_transform(ctx_, JGM_TF_NONE, 0.0, 0.0));
    return S;
  }

  void RawData(const rvec& m, size_t S) {
    raw_.assign(S * RawDataPerMeasurement(), 0.0);
    if (S > 0) Check(jgm_forward_host(ctx_, kind_, prec_, Ptr(m), raw_.data()));
  }

  // D(v)^T r for the calculator's transform D at raw data v, from the base class's
  // LQDerivative on the one-site probe: rows = identity, so its result is 2 D^T r exactly
  std::vector<double> TransformAdjoint(const double* v, const rvec& misfit, size_t off) {
    const size_t k = RawDataPerMeasurement(), kt = this->GetDataPerMeasurement();
    probe_raw_.assign(v, v + k);
    rvec r(kt);
    for (size_t j = 0; j < kt; ++j) r(j) = misfit(off + j);
    Probe guard(this);
    const rvec g = Base::LQDerivative(probe_model_, probe_data_, r);
    std::vector<double> a(k);
    for (size_t j = 0; j < k; ++j) a[j] = 0.5 * g(j);
    return a;
  }

  // whether the calculator set a transform: the output size differs from the raw size, or
  // the probe's Jacobian at a test vector is not the identity
  bool HasTransform() {
    const size_t k = RawDataPerMeasurement();
    if (this->GetDataPerMeasurement() != k) return true;
    std::vector<double> v(k);
    rvec r(k);
    for (size_t j = 0; j < k; ++j) {
      v[j] = 1.0 + 0.375 * static_cast<double>(j);
      r(j) = 2.0 - 0.25 * static_cast<double>(j);
    }
    const std::vector<double> a = TransformAdjoint(v.data(), r, 0);
    for (size_t j = 0; j < k; ++j)
      if (a[j] != r(j)) return true;
    return false;
  }

  // the FP64 rows of site i into the k x nmod sensitivity matrix, from blocks of sites
  // fetched from the GPU (at most ~64 MB per block)
  void FillRows(size_t i, rmat& sens) {
    const size_t k = RawDataPerMeasurement(), S = raw_.size() / k;
    if (i < row0_ || i >= row0_ + nrow_sites_) {
      const size_t per_site = k * nmod_ * sizeof(double);
      const size_t n = std::min(S - i, std::max<size_t>(1, (64u << 20) / per_site));
      const size_t bytes = n * per_site;
      if (bytes > rows_dev_bytes_) {
        if (rows_dev_) Check(jgm_device_free(ctx_, rows_dev_));
        rows_dev_ = nullptr;
        rows_dev_bytes_ = 0;
        Check(jgm_device_alloc(ctx_, bytes, &rows_dev_));
        rows_dev_bytes_ = bytes;
      }
      Check(jgm_sensitivity_rows(ctx_, kind_, i, n, static_cast<double*>(rows_dev_)));
      rows_.resize(n * k * nmod_);
      Check(jgm_copy_to_host(ctx_, rows_.data(), rows_dev_, bytes));
      row0_ = i;
      nrow_sites_ = n;
    }
    const double* src = rows_.data() + (i - row0_) * k * nmod_;
    for (size_t j = 0; j < k; ++j)
      for (size_t c = 0; c < nmod_; ++c) sens(j, c) = src[j * nmod_ + c];
  }

  // scope of the base class's site loop over the batch (rows invalidated on exit)
  struct Serving {
    HipGravMagImp* p;
    Serving(HipGravMagImp* imp, size_t nmod) : p(imp) {
      p->serving_ = true;
      p->nmod_ = nmod;
      p->row0_ = p->nrow_sites_ = 0;
    }
    ~Serving() {
      p->serving_ = false;
      p->nrow_sites_ = 0;
      std::vector<double>().swap(p->rows_);
    }
  };
  struct Probe {
    HipGravMagImp* p;
    explicit Probe(HipGravMagImp* imp) : p(imp) { p->probe_ = true; }
    ~Probe() { p->probe_ = false; }
  };

  jgm_context* ctx_ = nullptr;
  jgm_kind kind_;
  jgm_precision prec_;
  std::vector<double> geometry_;
  std::vector<double> raw_;       // raw data of every site (gridded + background)
  std::vector<double> rows_;      // FP64 rows of sites [row0_, row0_ + nrow_sites_)
  size_t row0_ = 0, nrow_sites_ = 0, nmod_ = 0;
  void* rows_dev_ = nullptr;
  size_t rows_dev_bytes_ = 0;
  bool serving_ = false, probe_ = false;
  std::vector<double> probe_raw_;
  ModelType probe_model_;
  DataType probe_data_;
};

}  // namespace jif3D
