// C++ host adapter over the C ABI of include/jif3d_gravmag.h.
//
// HipGravMagImplementation mirrors jif3D's ThreeDGravMagImplementation
// (jif3D/GravMag/ThreeDGravMagImplementation.h:39-84): RawDataPerMeasurement,
// SetDataTransform, GetDataPerMeasurement, Calculate and LQDerivative, with std::vector
// in place of ublas and a plain mesh/site description in place of the jif3D model and
// data classes.  Errors from the library are thrown as FatalException, as
// jif3D::FatalException is (jif3D/Global/FatalException.h:27-41).  The object owns one
// library context (device buffers, cached geometry); it is movable but not copyable,
// like the reference's CUDA imps (jif3D/Gravity/ScalarCudaGravityImp.h:75-76).
#pragma once

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "jif3d_gravmag.h"

namespace jif3d_amd {

class FatalException : public std::runtime_error {
 public:
  explicit FatalException(const std::string& msg) : std::runtime_error(msg) {}
};

inline void check(jgm_status st, const char* what) {
  if (st != JGM_OK)
    throw FatalException(std::string(what) + ": " + jgm_last_error() + " (status " +
                         std::to_string(static_cast<int>(st)) + ")");
}

// Rectilinear mesh as ThreeDModelBase stores it: n+1 corner coordinates and n cell
// sizes per axis, plus the 1-D background layer thicknesses.
struct Mesh {
  std::vector<double> xcoords, ycoords, zcoords;
  std::vector<double> xsizes, ysizes, zsizes;
  std::vector<double> bg_thicknesses;
  size_t ncells() const { return xsizes.size() * ysizes.size() * zsizes.size(); }
  bool operator==(const Mesh& o) const {
    return xcoords == o.xcoords && ycoords == o.ycoords && zcoords == o.zcoords &&
           xsizes == o.xsizes && ysizes == o.ysizes && zsizes == o.zsizes &&
           bg_thicknesses == o.bg_thicknesses;
  }
};

struct Sites {
  std::vector<double> x, y, z;
  size_t size() const { return x.size(); }
  bool operator==(const Sites& o) const { return x == o.x && y == o.y && z == o.z; }
};

class HipGravMagImplementation {
 public:
  HipGravMagImplementation(jgm_kind kind, jgm_precision prec = JGM_FP64, int device = 0,
                           void* stream = nullptr)
      : kind_(kind), prec_(prec) {
    if (jgm_api_version() != JGM_API_VERSION)
      throw FatalException("libjif3d_amd.so has C ABI version " +
                           std::to_string(jgm_api_version()) + ", this header " +
                           std::to_string(JGM_API_VERSION));
    check(jgm_create(device, stream, &ctx_), "jgm_create");
  }
  ~HipGravMagImplementation() {
    if (ctx_) jgm_destroy(ctx_);
  }
  HipGravMagImplementation(const HipGravMagImplementation&) = delete;
  HipGravMagImplementation& operator=(const HipGravMagImplementation&) = delete;
  HipGravMagImplementation(HipGravMagImplementation&& o) noexcept { *this = std::move(o); }
  HipGravMagImplementation& operator=(HipGravMagImplementation&& o) noexcept {
    std::swap(ctx_, o.ctx_);
    std::swap(kind_, o.kind_);
    std::swap(prec_, o.prec_);
    std::swap(mesh_, o.mesh_);
    std::swap(sites_, o.sites_);
    std::swap(have_geometry_, o.have_geometry_);
    return *this;
  }

  size_t RawDataPerMeasurement() const { return jgm_raw_per_site(kind_); }
  void SetDataTransform(jgm_transform tf, double inclination = 0.0, double declination = 0.0) {
    check(jgm_set_transform(ctx_, tf, inclination, declination), "jgm_set_transform");
    tf_ = tf;
  }
  size_t GetDataPerMeasurement() const { return tf_ == JGM_TF_NONE ? RawDataPerMeasurement() : 1; }
  // Inducing field of the susceptibility imp: radians, radians, nT.
  void SetField(double inclination, double declination, double fieldstrength) {
    check(jgm_set_field(ctx_, inclination, declination, fieldstrength), "jgm_set_field");
  }
  void SetPrecision(jgm_precision prec) { prec_ = prec; }

  // Forward response of every site (nsites * GetDataPerMeasurement values).
  std::vector<double> Calculate(const Mesh& mesh, const Sites& sites,
                                const std::vector<double>& model) {
    Configure(mesh, sites);
    if (model.size() != ModelSize())
      throw FatalException("model vector length does not match the mesh and background");
    std::vector<double> out(sites.size() * GetDataPerMeasurement());
    check(jgm_forward_host(ctx_, kind_, prec_, model.data(), out.data()), "jgm_forward_host");
    return out;
  }

  // 2 J^T D^T misfit, J recomputed (never stored).
  std::vector<double> LQDerivative(const Mesh& mesh, const Sites& sites,
                                   const std::vector<double>& model,
                                   const std::vector<double>& misfit) {
    Configure(mesh, sites);
    if (misfit.size() != sites.size() * GetDataPerMeasurement())
      throw FatalException("misfit length does not match the number of data");
    std::vector<double> grad(ModelSize());
    check(jgm_adjoint_host(ctx_, kind_, prec_, model.data(), misfit.data(), grad.data()),
          "jgm_adjoint_host");
    return grad;
  }

  // Calculate for nmodels model vectors at once (models: nmodels * ModelSize values,
  // model-major), e.g. the trial steps of a line search; returns nmodels * nsites *
  // GetDataPerMeasurement values.  FP64 batches share the corner-function evaluations.
  std::vector<double> CalculateBatch(const Mesh& mesh, const Sites& sites,
                                     const std::vector<double>& models, size_t nmodels) {
    Configure(mesh, sites);
    if (models.size() != nmodels * ModelSize())
      throw FatalException("models do not hold nmodels model vectors");
    std::vector<double> out(nmodels * sites.size() * GetDataPerMeasurement());
    check(jgm_forward_batch_host(ctx_, kind_, prec_, nmodels, models.data(), out.data()),
          "jgm_forward_batch_host");
    return out;
  }

  // LQDerivative for nmodels (model, misfit) pairs; models may be empty when no
  // transform is set.  Returns nmodels * ModelSize values.
  std::vector<double> LQDerivativeBatch(const Mesh& mesh, const Sites& sites,
                                        const std::vector<double>& models,
                                        const std::vector<double>& misfits, size_t nmodels) {
    Configure(mesh, sites);
    if (misfits.size() != nmodels * sites.size() * GetDataPerMeasurement())
      throw FatalException("misfits do not hold nmodels misfit vectors");
    if (!models.empty() && models.size() != nmodels * ModelSize())
      throw FatalException("models do not hold nmodels model vectors");
    std::vector<double> grads(nmodels * ModelSize());
    check(jgm_adjoint_batch_host(ctx_, kind_, prec_, nmodels,
                                 models.empty() ? nullptr : models.data(), misfits.data(),
                                 grads.data()),
          "jgm_adjoint_batch_host");
    return grads;
  }

  size_t ModelSize() const { return jgm_model_size(ctx_, kind_); }

  jgm_context* context() { return ctx_; }
  jgm_kind kind() const { return kind_; }
  jgm_precision precision() const { return prec_; }

  // Sets mesh, background and sites unless they are unchanged; returns true when the
  // geometry was (re)loaded.
  bool Configure(const Mesh& mesh, const Sites& sites) {
    if (have_geometry_ && mesh == mesh_ && sites == sites_) return false;
    check(jgm_set_mesh(ctx_, mesh.xsizes.size(), mesh.ysizes.size(), mesh.zsizes.size(),
                       mesh.xcoords.data(), mesh.ycoords.data(), mesh.zcoords.data(),
                       mesh.xsizes.data(), mesh.ysizes.data(), mesh.zsizes.data()),
          "jgm_set_mesh");
    check(jgm_set_background(ctx_, mesh.bg_thicknesses.size(), mesh.bg_thicknesses.data()),
          "jgm_set_background");
    check(jgm_set_sites(ctx_, sites.size(), sites.x.data(), sites.y.data(), sites.z.data()),
          "jgm_set_sites");
    mesh_ = mesh;
    sites_ = sites;
    have_geometry_ = true;
    return true;
  }

 private:
  jgm_context* ctx_ = nullptr;
  jgm_kind kind_ = JGM_KIND_SCALAR;
  jgm_precision prec_ = JGM_FP64;
  jgm_transform tf_ = JGM_TF_NONE;
  Mesh mesh_;
  Sites sites_;
  bool have_geometry_ = false;
};

// FullSensitivityGravMagCalculator with the geometry cache of CachedGravMagCalculator
// (jif3D/GravMag/FullSensitivityGravMagCalculator.h:99-221,
// CachedGravMagCalculator.h:27-135), J kept in HBM as FP32: the first Calculate for
// a geometry returns the imp's direct result and builds J; while mesh, background,
// sites and field are unchanged, Calculate is J m (plus the background layers) and
// LQDerivative is 2 [J B]^T D^T misfit on the matrix cores.
class HipFullSensitivityCalculator {
 public:
  explicit HipFullSensitivityCalculator(HipGravMagImplementation& imp,
                                        jgm_gemm_precision prec = JGM_GEMM_FP32)
      : imp_(imp), prec_(prec) {}

  std::vector<double> Calculate(const Mesh& mesh, const Sites& sites,
                                const std::vector<double>& model) {
    if (Cached(mesh, sites)) {
      std::vector<double> out(sites.size() * imp_.GetDataPerMeasurement());
      check(jgm_storedj_forward_host(imp_.context(), prec_, model.data(), out.data()),
            "jgm_storedj_forward_host");
      return out;
    }
    std::vector<double> out = imp_.Calculate(mesh, sites, model);
    check(jgm_storedj_build(imp_.context(), imp_.kind(), imp_.precision()), "jgm_storedj_build");
    return out;
  }

  std::vector<double> LQDerivative(const Mesh& mesh, const Sites& sites,
                                   const std::vector<double>& model,
                                   const std::vector<double>& misfit) {
    if (!Cached(mesh, sites)) return imp_.LQDerivative(mesh, sites, model, misfit);
    if (misfit.size() != sites.size() * imp_.GetDataPerMeasurement())
      throw FatalException("misfit length does not match the number of data");
    std::vector<double> grad(imp_.ModelSize());
    check(jgm_storedj_gradient_host(imp_.context(), prec_, model.data(), misfit.data(),
                                    grad.data()),
          "jgm_storedj_gradient_host");
    return grad;
  }

  // Batched forms: with a valid stored J up to 64 models share each pass over J;
  // otherwise the first model is calculated directly (building J) and the rest from J.
  std::vector<double> CalculateBatch(const Mesh& mesh, const Sites& sites,
                                     const std::vector<double>& models, size_t nmodels) {
    const size_t np = imp_.ModelSize(), nd = sites.size() * imp_.GetDataPerMeasurement();
    if (models.size() != nmodels * np)
      throw FatalException("models do not hold nmodels model vectors");
    std::vector<double> out(nmodels * nd);
    size_t first = 0;
    if (nmodels && !Cached(mesh, sites)) {
      std::vector<double> m0(models.begin(), models.begin() + np);
      std::vector<double> d0 = Calculate(mesh, sites, m0);
      std::copy(d0.begin(), d0.end(), out.begin());
      first = 1;
    }
    if (nmodels > first)
      check(jgm_storedj_forward_batch_host(imp_.context(), prec_, nmodels - first,
                                           models.data() + first * np, out.data() + first * nd),
            "jgm_storedj_forward_batch_host");
    return out;
  }

  std::vector<double> LQDerivativeBatch(const Mesh& mesh, const Sites& sites,
                                        const std::vector<double>& models,
                                        const std::vector<double>& misfits, size_t nmodels) {
    if (!Cached(mesh, sites)) return imp_.LQDerivativeBatch(mesh, sites, models, misfits, nmodels);
    const size_t np = imp_.ModelSize();
    if (misfits.size() != nmodels * sites.size() * imp_.GetDataPerMeasurement())
      throw FatalException("misfits do not hold nmodels misfit vectors");
    std::vector<double> grads(nmodels * np);
    check(jgm_storedj_gradient_batch_host(imp_.context(), prec_, nmodels,
                                          models.empty() ? nullptr : models.data(),
                                          misfits.data(), grads.data()),
          "jgm_storedj_gradient_batch_host");
    return grads;
  }

 private:
  // The stored J is valid while the geometry is unchanged; the library drops it itself
  // when the geometry or the inducing field changes (jgm_storedj_rows is then 0).
  bool Cached(const Mesh& mesh, const Sites& sites) {
    const bool reloaded = imp_.Configure(mesh, sites);
    return !reloaded && jgm_storedj_rows(imp_.context()) != 0;
  }

  HipGravMagImplementation& imp_;
  jgm_gemm_precision prec_;
};

}  // namespace jif3d_amd
