// The jif3D GravMag bridge classes restated on std::vector, so that an imp written for
// the jif3D tree compiles and runs outside it (the real headers need Boost, uBLAS and
// netCDF-C++).  Names, members, access levels and the order of calls follow the
// reference, because the imp in hip_bridge_imp.hpp is tested against exactly these
// contracts:
//
//   ThreeDGravMagImplementation   jif3D/GravMag/ThreeDGravMagImplementation.h:39-231
//     (private Transform and non-virtual SetDataTransform, :46, :67-70; the site loops of
//     Calculate, :116-168, and LQDerivative, :177-231)
//   ThreeDGravMagCalculator        jif3D/GravMag/ThreeDGravMagCalculator.h:60-173
//   MinMemGravMagCalculator        jif3D/GravMag/MinMemGravMagCalculator.h:60-79
//   CachedGravMagCalculator +
//   FullSensitivityGravMagCalculator  jif3D/GravMag/CachedGravMagCalculator.h:98-143,
//                                  jif3D/GravMag/FullSensitivityGravMagCalculator.h:99-221
//   FTGInvariant                   jif3D/Gravity/GravityTransforms.h:25-100
//   TotalField, TotalFieldAnomaly  jif3D/Magnetics/MagneticTransforms.h:19-156
//
// rvec / rmat stand for uBLAS vector<double> / matrix<double> (row-major), with the
// members the bridge uses.  The model class keeps what ThreeDModelBase and the gravity /
// susceptibility models expose to an imp (cell sizes, n+1 corner coordinates per axis,
// background layers, model vector); the data classes keep the measurement positions.
#pragma once

#include <cmath>
#include <cstddef>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace jif3D {

class FatalException : public std::runtime_error {
 public:
  explicit FatalException(const std::string& msg, const char* file = "", int line = 0)
      : std::runtime_error(msg + (file && *file ? std::string(" (") + file + ":" +
                                                      std::to_string(line) + ")"
                                                : std::string())) {}
};

class rvec {
 public:
  explicit rvec(size_t n = 0) : v_(n, 0.0) {}
  size_t size() const { return v_.size(); }
  double& operator()(size_t i) { return v_[i]; }
  const double& operator()(size_t i) const { return v_[i]; }
  double* begin() { return v_.data(); }
  double* end() { return v_.data() + v_.size(); }
  const double* begin() const { return v_.data(); }
  const double* end() const { return v_.data() + v_.size(); }
  void resize(size_t n) { v_.resize(n, 0.0); }
  rvec& operator+=(const rvec& o) {
    for (size_t i = 0; i < v_.size(); ++i) v_[i] += o.v_[i];
    return *this;
  }

 private:
  std::vector<double> v_;
};

inline rvec operator*(double a, const rvec& x) {
  rvec r(x.size());
  for (size_t i = 0; i < x.size(); ++i) r(i) = a * x(i);
  return r;
}

class rmat {
 public:
  rmat(size_t r = 0, size_t c = 0) : r_(r), c_(c), v_(r * c, 0.0) {}
  size_t size1() const { return r_; }
  size_t size2() const { return c_; }
  double& operator()(size_t i, size_t j) { return v_[i * c_ + j]; }
  const double& operator()(size_t i, size_t j) const { return v_[i * c_ + j]; }
  void resize(size_t r, size_t c, bool = true) {
    r_ = r;
    c_ = c;
    v_.assign(r * c, 0.0);
  }
  void clear() { std::fill(v_.begin(), v_.end(), 0.0); }

 private:
  size_t r_, c_;
  std::vector<double> v_;
};

// ---- data transforms ------------------------------------------------------------------
class VectorTransform {
 public:
  virtual ~VectorTransform() = default;
  virtual size_t GetInputSize() const = 0;
  virtual size_t GetOutputSize() const = 0;
  virtual rvec Transform(const rvec& in) const = 0;
  virtual rmat Derivative(const rvec& in) const = 0;
};

// the invariant Uxx Uyy + Uyy Uzz + Uxx Uzz - Uxy^2 - Uyz^2 - Uxz^2 of the 9 row-major
// tensor values (the off-diagonal pairs taken as the products of both entries)
class FTGInvariant : public VectorTransform {
 public:
  size_t GetInputSize() const override { return 9; }
  size_t GetOutputSize() const override { return 1; }
  rvec Transform(const rvec& u) const override {
    rvec r(1);
    r(0) = u(0) * u(4) + u(4) * u(8) + u(0) * u(8) - u(3) * u(1) - u(7) * u(5) - u(2) * u(6);
    return r;
  }
  rmat Derivative(const rvec& u) const override {
    rmat d(1, 9);
    d(0, 0) = u(4) + u(8);
    d(0, 1) = -u(3);
    d(0, 2) = -u(6);
    d(0, 3) = -u(1);
    d(0, 4) = u(0) + u(8);
    d(0, 5) = -u(7);
    d(0, 6) = -u(2);
    d(0, 7) = -u(5);
    d(0, 8) = u(0) + u(4);
    return d;
  }
};

class TotalField : public VectorTransform {
 public:
  size_t GetInputSize() const override { return 3; }
  size_t GetOutputSize() const override { return 1; }
  rvec Transform(const rvec& b) const override {
    rvec r(1);
    r(0) = std::sqrt(b(0) * b(0) + b(1) * b(1) + b(2) * b(2));
    return r;
  }
  rmat Derivative(const rvec& b) const override {
    const double t = std::sqrt(b(0) * b(0) + b(1) * b(1) + b(2) * b(2));
    rmat d(1, 3);
    for (size_t i = 0; i < 3; ++i) d(0, i) = b(i) / t;
    return d;
  }
};

class TotalFieldAnomaly : public VectorTransform {
 public:
  explicit TotalFieldAnomaly(double inc = 0.0, double dec = 0.0)
      : bx_(std::cos(inc) * std::cos(dec)), by_(std::cos(inc) * std::sin(dec)),
        bz_(std::sin(inc)) {}
  size_t GetInputSize() const override { return 3; }
  size_t GetOutputSize() const override { return 1; }
  rvec Transform(const rvec& b) const override {
    rvec r(1);
    r(0) = bx_ * b(0) + by_ * b(1) + bz_ * b(2);
    return r;
  }
  rmat Derivative(const rvec&) const override {
    rmat d(1, 3);
    d(0, 0) = bx_;
    d(0, 1) = by_;
    d(0, 2) = bz_;
    return d;
  }

 private:
  double bx_, by_, bz_;
};

// ---- model and data -------------------------------------------------------------------
class GravMagModel {
 public:
  typedef std::vector<double> t3DModelDim;
  virtual ~GravMagModel() = default;
  virtual void SetMeshSize(size_t nx, size_t ny, size_t nz) {
    xs_.assign(nx, 0.0);
    ys_.assign(ny, 0.0);
    zs_.assign(nz, 0.0);
    Recalc();
    values_.assign(nx * ny * nz, 0.0);
  }
  void SetOrigin(double x, double y, double z) {
    origin_[0] = x;
    origin_[1] = y;
    origin_[2] = z;
    Recalc();
  }
  void SetXCellSizes(const t3DModelDim& s) { xs_ = s; Recalc(); }
  void SetYCellSizes(const t3DModelDim& s) { ys_ = s; Recalc(); }
  void SetZCellSizes(const t3DModelDim& s) { zs_ = s; Recalc(); }
  const t3DModelDim& GetXCellSizes() const { return xs_; }
  const t3DModelDim& GetYCellSizes() const { return ys_; }
  const t3DModelDim& GetZCellSizes() const { return zs_; }
  const t3DModelDim& GetXCoordinates() const { return xc_; }
  const t3DModelDim& GetYCoordinates() const { return yc_; }
  const t3DModelDim& GetZCoordinates() const { return zc_; }
  void SetBackgroundThicknesses(const std::vector<double>& t) { bg_thick_ = t; }
  void SetBackgroundValues(const std::vector<double>& v) { bg_values_ = v; }
  const std::vector<double>& GetBackgroundThicknesses() const { return bg_thick_; }
  std::vector<double>& SetValues() { return values_; }
  rvec GetModelParameters() const {
    rvec p(values_.size() + bg_values_.size());
    std::copy(values_.begin(), values_.end(), p.begin());
    std::copy(bg_values_.begin(), bg_values_.end(), p.begin() + values_.size());
    return p;
  }
  size_t GetNModelParm() const { return values_.size() + bg_values_.size(); }

 private:
  // n+1 corner coordinates from the origin and the sizes (ThreeDModelBase::CalcCoords)
  void Recalc() {
    const t3DModelDim* s[3] = {&xs_, &ys_, &zs_};
    t3DModelDim* c[3] = {&xc_, &yc_, &zc_};
    for (int a = 0; a < 3; ++a) {
      c[a]->assign(s[a]->size() + 1, origin_[a]);
      for (size_t i = 0; i < s[a]->size(); ++i) (*c[a])[i + 1] = (*c[a])[i] + (*s[a])[i];
    }
  }
  t3DModelDim xs_, ys_, zs_, xc_, yc_, zc_;
  double origin_[3] = {0.0, 0.0, 0.0};
  std::vector<double> values_, bg_thick_, bg_values_;
};

typedef GravMagModel ThreeDGravityModel;
typedef GravMagModel ThreeDSusceptibilityModel;

class GeneralData {
 public:
  typedef std::vector<double> tMeasPosVec;
  void AddMeasurementPoint(double x, double y, double z) {
    x_.push_back(x);
    y_.push_back(y);
    z_.push_back(z);
  }
  const tMeasPosVec& GetMeasPosX() const { return x_; }
  const tMeasPosVec& GetMeasPosY() const { return y_; }
  const tMeasPosVec& GetMeasPosZ() const { return z_; }

 private:
  tMeasPosVec x_, y_, z_;
};

struct ScalarGravityData : GeneralData { typedef ThreeDGravityModel ModelType; };
struct TensorGravityData : GeneralData { typedef ThreeDGravityModel ModelType; };
struct TotalFieldMagneticData : GeneralData { typedef ThreeDSusceptibilityModel ModelType; };

// ---- bridge ---------------------------------------------------------------------------
template <class PotentialDataType>
class ThreeDGravMagCalculator;

template <class PotentialDataType>
class ThreeDGravMagImplementation {
 public:
  typedef typename PotentialDataType::ModelType ThreeDModelType;

 private:
  std::shared_ptr<VectorTransform> Transform;
  virtual rvec CalcBackground(const size_t measindex, const double xwidth, const double ywidth,
                              const double zwidth, const ThreeDModelType& Model,
                              const PotentialDataType& Data, rmat& Sensitivities) = 0;
  virtual rvec CalcGridded(const size_t measindex, const ThreeDModelType& Model,
                           const PotentialDataType& Data, rmat& Sensitivities) = 0;

 public:
  virtual size_t RawDataPerMeasurement() = 0;
  void SetDataTransform(std::shared_ptr<VectorTransform> DataTransform) {
    Transform = DataTransform;
  }
  size_t GetDataPerMeasurement() {
    return Transform ? Transform->GetOutputSize() : RawDataPerMeasurement();
  }
  virtual rvec Calculate(const ThreeDModelType& Model, const PotentialDataType& Data,
                         ThreeDGravMagCalculator<PotentialDataType>& Calculator);
  virtual rvec LQDerivative(const ThreeDModelType& Model, const PotentialDataType& Data,
                            const rvec& Misfit);
  virtual ~ThreeDGravMagImplementation() = default;
};

template <class D>
rvec ThreeDGravMagImplementation<D>::Calculate(const ThreeDModelType& Model, const D& Data,
                                               ThreeDGravMagCalculator<D>& Calculator) {
  const size_t nmeas = Data.GetMeasPosX().size();
  const size_t nout = GetDataPerMeasurement();
  rvec result(nmeas * nout);
  const auto& xs = Model.GetXCellSizes();
  const auto& ys = Model.GetYCellSizes();
  const auto& zs = Model.GetZCellSizes();
  const double xw = std::accumulate(xs.begin(), xs.end(), 0.0);
  const double yw = std::accumulate(ys.begin(), ys.end(), 0.0);
  const double zw = std::accumulate(zs.begin(), zs.end(), 0.0);
  for (size_t i = 0; i < nmeas; ++i) {
    rvec currdata = CalcGridded(i, Model, Data, Calculator.SetCurrentSensitivities());
    currdata += CalcBackground(i, xw, yw, zw, Model, Data, Calculator.SetCurrentSensitivities());
    const rvec out = Transform ? Transform->Transform(currdata) : currdata;
    std::copy(out.begin(), out.end(), result.begin() + i * nout);
    Calculator.HandleSensitivities(i);
  }
  return result;
}

template <class D>
rvec ThreeDGravMagImplementation<D>::LQDerivative(const ThreeDModelType& Model, const D& Data,
                                                  const rvec& Misfit) {
  const size_t nmeas = Data.GetMeasPosX().size();
  const size_t nout = GetDataPerMeasurement();
  const size_t k = RawDataPerMeasurement();
  if (Misfit.size() != nmeas * nout) throw FatalException("misfit size", __FILE__, __LINE__);
  const auto& xs = Model.GetXCellSizes();
  const auto& ys = Model.GetYCellSizes();
  const auto& zs = Model.GetZCellSizes();
  const double xw = std::accumulate(xs.begin(), xs.end(), 0.0);
  const double yw = std::accumulate(ys.begin(), ys.end(), 0.0);
  const double zw = std::accumulate(zs.begin(), zs.end(), 0.0);
  const size_t nmod = Model.GetNModelParm();
  rvec DerivMod(nmod);
  rmat Current(k, nmod);
  for (size_t i = 0; i < nmeas; ++i) {
    rvec currdata = CalcGridded(i, Model, Data, Current);
    currdata += CalcBackground(i, xw, yw, zw, Model, Data, Current);
    // NewSensitivities = TransDeriv x CurrentSensitivities
    rmat New(nout, nmod);
    if (Transform) {
      const rmat T = Transform->Derivative(currdata);
      for (size_t r = 0; r < nout; ++r)
        for (size_t c = 0; c < nmod; ++c) {
          double s = 0.0;
          for (size_t j = 0; j < k; ++j) s += T(r, j) * Current(j, c);
          New(r, c) = s;
        }
    } else {
      New = Current;
    }
    for (size_t j = 0; j < nout; ++j)
      for (size_t c = 0; c < nmod; ++c) DerivMod(c) += Misfit(i * nout + j) * New(j, c);
  }
  return 2.0 * DerivMod;
}

template <class PotentialDataType>
class ThreeDGravMagCalculator {
 public:
  typedef PotentialDataType DataType;
  typedef typename PotentialDataType::ModelType ThreeDModelType;
  explicit ThreeDGravMagCalculator(
      std::shared_ptr<ThreeDGravMagImplementation<PotentialDataType>> TheImp)
      : Imp(TheImp) {}
  virtual ~ThreeDGravMagCalculator() = default;
  void SetDataTransform(std::shared_ptr<VectorTransform> DataTransform) {
    Transform = DataTransform;
    Imp->SetDataTransform(DataTransform);
  }
  virtual rvec Calculate(const ThreeDModelType& Model, const PotentialDataType& Data) {
    return Imp->Calculate(Model, Data, *this);
  }
  virtual rvec LQDerivative(const ThreeDModelType& Model, const PotentialDataType& Data,
                            const rvec& Misfit) {
    return Imp->LQDerivative(Model, Data, Misfit);
  }
  rmat& SetCurrentSensitivities() { return CurrentSensitivities; }
  size_t GetDataPerMeasurement() { return Imp->GetDataPerMeasurement(); }
  virtual void HandleSensitivities(const size_t measindex) = 0;

 private:
  rmat CurrentSensitivities;

 protected:
  std::shared_ptr<VectorTransform> Transform;
  std::shared_ptr<ThreeDGravMagImplementation<PotentialDataType>> Imp;
};

template <class D>
class MinMemGravMagCalculator : public ThreeDGravMagCalculator<D> {
 public:
  using ThreeDGravMagCalculator<D>::ThreeDGravMagCalculator;
  rvec Calculate(const typename D::ModelType& Model, const D& Data) override {
    this->SetCurrentSensitivities().resize(0, 0);
    return ThreeDGravMagCalculator<D>::Calculate(Model, Data);
  }
  void HandleSensitivities(const size_t) override {}
};

// CachedGravMagCalculator's decision (a new geometry recomputes and stores J, the same
// geometry takes the cached path) and FullSensitivityGravMagCalculator's J in one class.
// The geometry key is mesh, background layers and sites.
template <class D>
class FullSensitivityGravMagCalculator : public ThreeDGravMagCalculator<D> {
 public:
  using ThreeDGravMagCalculator<D>::ThreeDGravMagCalculator;
  rvec Calculate(const typename D::ModelType& Model, const D& Data) override {
    const bool changed = GeometryChanged(Model, Data);
    if (HaveCache && !changed) return CalculateCachedResult(Model);
    this->SetCurrentSensitivities().resize(this->Imp->RawDataPerMeasurement(),
                                           Model.GetNModelParm());
    const size_t nmeas = Data.GetMeasPosX().size();
    Sensitivities.resize(nmeas * this->Imp->RawDataPerMeasurement(), Model.GetNModelParm());
    rvec result = this->Imp->Calculate(Model, Data, *this);
    HaveCache = true;
    return result;
  }
  rvec LQDerivative(const typename D::ModelType& Model, const D& Data,
                    const rvec& Misfit) override {
    const bool changed = GeometryChanged(Model, Data);
    if (HaveCache && !changed) return CachedLQDerivative(Model, Misfit);
    HaveCache = false;  // J belongs to another geometry
    this->SetCurrentSensitivities().resize(this->Imp->GetDataPerMeasurement(),
                                           Model.GetNModelParm());
    return ThreeDGravMagCalculator<D>::LQDerivative(Model, Data, Misfit);
  }
  void HandleSensitivities(const size_t measindex) override {
    const size_t k = this->Imp->RawDataPerMeasurement();
    const rmat& cur = this->SetCurrentSensitivities();
    for (size_t j = 0; j < k; ++j)
      for (size_t c = 0; c < Sensitivities.size2(); ++c)
        Sensitivities(measindex * k + j, c) = cur(j, c);
  }
  const rmat& GetSensitivities() const { return Sensitivities; }

 private:
  rvec RawData(const typename D::ModelType& Model) const {
    const rvec m = Model.GetModelParameters();
    rvec d(Sensitivities.size1());
    for (size_t r = 0; r < Sensitivities.size1(); ++r) {
      double s = 0.0;
      for (size_t c = 0; c < Sensitivities.size2(); ++c) s += Sensitivities(r, c) * m(c);
      d(r) = s;
    }
    return d;
  }
  rvec CalculateCachedResult(const typename D::ModelType& Model) const {
    rvec raw = RawData(Model);
    if (!this->Transform) return raw;
    const size_t nin = this->Transform->GetInputSize(), nout = this->Transform->GetOutputSize();
    rvec out(raw.size() / nin * nout);
    for (size_t i = 0; i < raw.size() / nin; ++i) {
      rvec seg(nin);
      for (size_t j = 0; j < nin; ++j) seg(j) = raw(i * nin + j);
      const rvec t = this->Transform->Transform(seg);
      for (size_t j = 0; j < nout; ++j) out(i * nout + j) = t(j);
    }
    return out;
  }
  rvec CachedLQDerivative(const typename D::ModelType& Model, const rvec& Misfit) const {
    rvec processed = Misfit;
    if (this->Transform) {
      const rvec raw = RawData(Model);
      const size_t nin = this->Transform->GetInputSize(), nout = this->Transform->GetOutputSize();
      processed = rvec(Misfit.size() / nout * nin);
      for (size_t i = 0; i < Misfit.size() / nout; ++i) {
        rvec seg(nin);
        for (size_t j = 0; j < nin; ++j) seg(j) = raw(i * nin + j);
        const rmat T = this->Transform->Derivative(seg);
        for (size_t j = 0; j < nin; ++j) {
          double s = 0.0;
          for (size_t o = 0; o < nout; ++o) s += T(o, j) * Misfit(i * nout + o);
          processed(i * nin + j) = s;
        }
      }
    }
    rvec g(Sensitivities.size2());
    for (size_t c = 0; c < Sensitivities.size2(); ++c) {
      double s = 0.0;
      for (size_t r = 0; r < Sensitivities.size1(); ++r) s += Sensitivities(r, c) * processed(r);
      g(c) = 2.0 * s;
    }
    return g;
  }
  bool GeometryChanged(const typename D::ModelType& M, const D& Data) {
    std::vector<double> key;
    for (const auto* v : {&M.GetXCoordinates(), &M.GetYCoordinates(), &M.GetZCoordinates(),
                          &M.GetBackgroundThicknesses(), &Data.GetMeasPosX(), &Data.GetMeasPosY(),
                          &Data.GetMeasPosZ()}) {
      key.push_back((double)v->size());
      key.insert(key.end(), v->begin(), v->end());
    }
    const bool changed = key != geometry_;
    geometry_ = key;
    return changed;
  }
  rmat Sensitivities;
  bool HaveCache = false;
  std::vector<double> geometry_;
};

}  // namespace jif3D
