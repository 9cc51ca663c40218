/*
 * jif3d_gravmag.h -- C ABI of the MI355X prism gravity / FTG / magnetic library.
 *
 * One shared library (libjif3d_amd.so) replaces the compute behind jif3D's
 * ThreeDGravMagImplementation (jif3D/GravMag/ThreeDGravMagImplementation.h:39-84)
 * for the scalar, tensor and magnetic-susceptibility imps and the stored-J
 * products of FullSensitivityGravMagCalculator.  Plain C types only; cell values
 * are z-fastest (offset = (xi*ny + yi)*nz + zi, jif3D/ModelBase/ThreeDModelBase.h:157-168);
 * SI units (kg/m^3, m, nT, radians); z positive down.
 *
 * Pointer conventions: geometry setters take host arrays and copy them.  The
 * plain compute calls take device pointers on the context's device and run on
 * the context's stream (asynchronously); the *_host variants take host arrays
 * and return after the result is copied back.
 *
 * Errors: every call returns a jgm_status; jgm_last_error() gives the message of
 * the last failure on the calling thread.  Nothing calls exit() and nothing resets
 * the device (compare jif3D/Global/CudaError.h:13-22 and the cudaDeviceReset in
 * jif3D/Gravity/ScalarCudaGravityImp.cpp:52).
 *
 * Threading: a context is used by one thread at a time; distinct contexts are
 * independent.
 */
#ifndef JIF3D_GRAVMAG_H_
#define JIF3D_GRAVMAG_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JGM_API_VERSION 1

typedef struct jgm_context jgm_context;

typedef enum jgm_status {
  JGM_OK = 0,
  JGM_ERR_INVALID_ARGUMENT = 1, /* sizes, null pointers, unknown enum values */
  JGM_ERR_NOT_READY = 2,        /* mesh or sites not set */
  JGM_ERR_DEVICE = 3,           /* HIP runtime error */
  JGM_ERR_OUT_OF_MEMORY = 4,
  JGM_ERR_UNSUPPORTED = 5
} jgm_status;

/* Raw data per site: 1 (gz, m/s^2), 9 (FTG, 1/s^2, row-major Uxx..Uzz as
 * jif3D/Gravity/TensorOMPGravityImp.cpp:255-264), 3 (Bx, By, Bz in nT), or 10
 * (gz followed by the 9 FTG values: one pass for a joint scalar+tensor survey).
 * JGM_KIND_MAGNETIC_GRADIENT: 1 value, Bz at the site minus Bz 1 m below with
 * mu0/(4 pi) = 1e-7 (OMPMagneticGradImp, jif3D/Magnetics/OMPMagneticGradImp.cpp:13-106;
 * no background).  JGM_KIND_MAGNETIZATION: Bx, By, Bz of a magnetization model with
 * three values per cell, [Mx; My; Mz], B = 1e9 mu0/(4 pi G) T M (OMPMagnetizationImp,
 * jif3D/Magnetics/OMPMagnetizationImp.cpp:25-117; no background, model length
 * 3 ncells; JGM_MIXED evaluates the tensor far field once per site-cell pair for the
 * three components).  JGM_MIXED evaluates the vertical gradient as the Bz of 1 m slabs
 * just above the mesh planes (the difference of the two fields, exactly), with FP64 for
 * sites on mesh lines; a mesh with a cell thinner than 1 m runs it in JGM_FP64. */
typedef enum jgm_kind {
  JGM_KIND_SCALAR = 1,
  JGM_KIND_TENSOR = 2,
  JGM_KIND_MAGNETIC = 3,
  JGM_KIND_SCALAR_TENSOR = 4,
  JGM_KIND_MAGNETIC_GRADIENT = 5,
  JGM_KIND_MAGNETIZATION = 6
} jgm_kind;

/* JGM_FP64: closed-form box terms in FP64 for every site-cell pair.
 * JGM_MIXED: FP64 closed form for cells inside each site's near box (every cell closer
 * than q0 of its own half-diagonals: the layers within their reach of the site, and along
 * x and y the largest reach of those layers), FP32 multipole of the prism about its centre
 * through the hexadecapole for the rest, FP64 accumulation across tiles. */
typedef enum jgm_precision { JGM_FP64 = 0, JGM_MIXED = 1 } jgm_precision;

/* Per-site data transforms (jif3D/Gravity/GravityTransforms.h:25-100,
 * jif3D/Magnetics/MagneticTransforms.h:19-156). */
typedef enum jgm_transform {
  JGM_TF_NONE = 0,
  JGM_TF_FTG_INVARIANT = 1,     /* tensor kind only */
  JGM_TF_TOTAL_FIELD = 2,       /* magnetic and magnetization kinds */
  JGM_TF_TOTAL_FIELD_ANOMALY = 3 /* magnetic and magnetization kinds */
} jgm_transform;

/* Precision of the stored-J products. */
typedef enum jgm_gemm_precision {
  JGM_GEMM_FP32 = 0,        /* FP32 MFMA, FP32 accumulate */
  JGM_GEMM_BF16 = 1,        /* BF16 operands, FP32 accumulate */
  JGM_GEMM_BF16_REFINED = 2, /* BF16 hi/lo split of both operands, three products */
  JGM_GEMM_FP64 = 3          /* FP64 J (jgm_storedj_build_fp64) on FP64 MFMA, FP64 accumulate */
} jgm_gemm_precision;

const char *jgm_last_error(void);
int jgm_api_version(void);
size_t jgm_raw_per_site(jgm_kind kind);

/* Create a context on HIP device `device`; `stream` is a hipStream_t or NULL for
 * the device's null stream.  Replaces PrepareData's device pick
 * (jif3D/Gravity/gravcuda.cu:38-107). */
jgm_status jgm_create(int device, void *stream, jgm_context **out);
jgm_status jgm_destroy(jgm_context *ctx);
jgm_status jgm_set_stream(jgm_context *ctx, void *stream);
jgm_status jgm_synchronize(jgm_context *ctx);

/* Mesh: n+1 corner coordinates and n cell sizes per axis, as
 * ThreeDModelBase::GetX/Y/ZCoordinates and Get*CellSizes return them; mirrors the
 * argument list of PrepareData (jif3D/Gravity/gravcuda.cu:38-44).  Cell i spans
 * [coord[i], coord[i] + size[i]].  Any of nx, ny, nz may be zero (background only). */
jgm_status jgm_set_mesh(jgm_context *ctx, size_t nx, size_t ny, size_t nz, const double *xcoords,
                        const double *ycoords, const double *zcoords, const double *xsizes,
                        const double *ysizes, const double *zsizes);
/* 1-D background layer thicknesses (ThreeDGravityModel::SetBackgroundThicknesses);
 * the layer values are the last nlayers entries of the model vector. */
jgm_status jgm_set_background(jgm_context *ctx, size_t nlayers, const double *thicknesses);
/* Measurement sites (GeneralData::GetMeasPosX/Y/Z, jif3D/DataBase/GeneralData.h:95-108). */
jgm_status jgm_set_sites(jgm_context *ctx, size_t nsites, const double *x, const double *y,
                         const double *z);
/* Inducing field of the susceptibility imp (OMPMagneticSusceptibilityImp(Inc, Dec, Fs),
 * jif3D/Magnetics/OMPMagneticSusceptibilityImp.cpp:196-201): radians, radians, nT. */
jgm_status jgm_set_field(jgm_context *ctx, double inclination, double declination,
                         double fieldstrength);
/* Data transform (ThreeDGravMagImplementation::SetDataTransform, :67-70). */
jgm_status jgm_set_transform(jgm_context *ctx, jgm_transform tf, double inclination,
                             double declination);
/* Near-field radius of JGM_MIXED in half-diagonals of each layer's largest cell shape
 * (default 8). */
jgm_status jgm_set_mixed_threshold(jgm_context *ctx, double q0);
/* Second radius of JGM_MIXED in cell half-diagonals (default 28): cells farther than
 * q1 from a site keep the monopole (and, for non-cubic cells, quadrupole) terms only;
 * the hexadecapole term dropped there is O((h/d)^4) of each cell's field.  0 keeps the
 * hexadecapole for every far cell. */
jgm_status jgm_set_mixed_far_radius(jgm_context *ctx, double q1);

/* A-posteriori check of JGM_MIXED (on by default: 4 sites, 16 cells, tolerance 5e-7).
 * The mixed path's error bound is relative to each component's largest magnitude over the
 * survey (forward) and to the largest gradient entry (adjoint); it holds unless the cells'
 * contributions cancel strongly (a nearly uniformly magnetised wide slab, whose external
 * field is a small remainder of its cells' fields).  After a mixed forward of the scalar,
 * tensor, magnetic or scalar+tensor kind the FP64 closed form is evaluated at nsites sample
 * sites spread over the survey, after a mixed adjoint at ncells sample cells spread over the
 * mesh, and when the largest sampled error, relative as above, exceeds `tolerance` the whole
 * call is recomputed in FP64.  The check waits for the context's stream once per call.
 * nsites = 0 turns it off (ncells = 0 turns off the adjoint part).  No counterpart in the
 * reference, whose OMP imps are FP64 throughout. */
jgm_status jgm_set_mixed_check(jgm_context *ctx, size_t nsites, size_t ncells, double tolerance);
/* The sampled errors of the last checked mixed forward and adjoint, and the number of calls
 * recomputed in FP64 since the context was created. */
jgm_status jgm_mixed_check_report(const jgm_context *ctx, double *forward_error,
                                  double *adjoint_error, size_t *fallbacks);

/* Values per site after the transform, and the model vector length
 * (ncells + nlayers, ThreeDGravityModel::GetNModelParm). */
size_t jgm_data_per_site(const jgm_context *ctx, jgm_kind kind);
size_t jgm_n_params(const jgm_context *ctx);
/* Model vector length of a kind: 3 ncells for JGM_KIND_MAGNETIZATION
 * (ThreeDMagnetizationModel), jgm_n_params otherwise. */
size_t jgm_model_size(const jgm_context *ctx, jgm_kind kind);

/* Forward response of every site: ThreeDGravMagImplementation::Calculate
 * (jif3D/GravMag/ThreeDGravMagImplementation.h:116-168) with the OMP imps'
 * CalcGridded/CalcBackground (jif3D/Gravity/ScalarOMPGravityImp.cpp:124-180,
 * TensorOMPGravityImp.cpp:112-182, OMPMagneticSusceptibilityImp.cpp:31-194).
 * model: jgm_model_size(kind) doubles (device); out: nsites * data_per_site doubles
 * (device). */
jgm_status jgm_forward(jgm_context *ctx, jgm_kind kind, jgm_precision prec, const double *model,
                       double *out);
jgm_status jgm_forward_host(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                            const double *model, double *out);

/* Allocation-free form of jgm_forward, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  It calls no hipMalloc / hipFree and no allocating
 * library routine: its working memory (background table, corner weights, partial sums)
 * is one static device array of JGM_NOALLOC_SCRATCH_BYTES per GPU, reserved when the
 * library is loaded and held until the process ends.  Same arguments and results as
 * jgm_forward with JGM_FP64 (sums grouped into fewer chunks, so results agree to FP64
 * rounding, not bit for bit); JGM_MIXED, JGM_KIND_MAGNETIC_GRADIENT and
 * JGM_KIND_MAGNETIZATION return JGM_ERR_UNSUPPORTED, and so does a problem whose
 * (mesh corners + 7 nsites nlayers + components nsites + nlayers) doubles exceed the
 * scratch.  Unlike jgm_forward it keeps no raw data for a later jgm_adjoint.  The mesh,
 * background and sites must have been set (those setters allocate).  All calls on one
 * GPU share its scratch array: calls that overlap in time on one GPU (other threads,
 * other streams, other contexts) race and must be serialised by the caller. */
#define JGM_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
jgm_status jgm_forward_noalloc(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                               const double *model, double *out);

/* Least-squares derivative 2 J^T D^T misfit, matrix free (J recomputed, never stored):
 * ThreeDGravMagImplementation::LQDerivative (:177-231).  model is needed for the
 * transform Jacobian D.  misfit: nsites * data_per_site; grad: jgm_model_size(kind)
 * doubles (device).
 * With a transform set, jgm_forward keeps its raw data and a copy of its model, and
 * jgm_adjoint uses them when its model is bitwise equal to that copy and no setting
 * has changed since (it then waits for the context's stream once, to read the
 * comparison); otherwise it computes the raw data again.  Results are the same
 * either way. */
jgm_status jgm_adjoint(jgm_context *ctx, jgm_kind kind, jgm_precision prec, const double *model,
                       const double *misfit, double *grad);
jgm_status jgm_adjoint_host(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                            const double *model, const double *misfit, double *grad);

/* Batched forms of jgm_forward / jgm_adjoint for several models at once, e.g. the
 * trial steps of a line search, which the reference evaluates one Calculate /
 * LQDerivative pair at a time (jif3D/Inversion/mcsrch.cpp:247-248 through
 * ObjectiveFunction::CalcMisfit/CalcGradient, ObjectiveFunction.h:165-229).
 * Model b is models[b * model_ld ...] (model_ld >= jgm_model_size), its data
 * out[b * out_ld ...] (out_ld >= nsites * data_per_site), its misfit
 * misfits[b * misfit_ld ...] and its gradient grads[b * grad_ld ...] (grad_ld >=
 * jgm_model_size); all device memory.  With JGM_FP64 and the gravity and
 * susceptibility kinds up to 8 models share each corner-function evaluation, so a
 * batch of 8 costs little more than one model; with JGM_MIXED up to 4 models share each
 * far-field multipole evaluation.  Each result is bit-identical to the single-model
 * call.  The vertical-gradient and magnetization kinds run the models one after
 * another (the magnetization kind's three components share the far-field terms within
 * each model).  nmodels == 0 is a no-op. */
jgm_status jgm_forward_batch(jgm_context *ctx, jgm_kind kind, jgm_precision prec, size_t nmodels,
                             const double *models, size_t model_ld, double *out, size_t out_ld);
jgm_status jgm_adjoint_batch(jgm_context *ctx, jgm_kind kind, jgm_precision prec, size_t nmodels,
                             const double *models, size_t model_ld, const double *misfits,
                             size_t misfit_ld, double *grads, size_t grad_ld);
/* The same with host memory and packed batches (model_ld = jgm_model_size, out_ld =
 * misfit_ld = nsites * data_per_site, grad_ld = jgm_model_size); models may be NULL
 * for jgm_adjoint_batch_host when no transform is set. */
jgm_status jgm_forward_batch_host(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                                  size_t nmodels, const double *models, double *out);
jgm_status jgm_adjoint_batch_host(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                                  size_t nmodels, const double *models, const double *misfits,
                                  double *grads);

/* Raw (untransformed) FP64 sensitivity rows of sites [site0, site0+nsites) in the
 * layout FullSensitivityGravMagCalculator::HandleSensitivities builds
 * (jif3D/GravMag/FullSensitivityGravMagCalculator.h:99-114): nsites*raw_per_site rows
 * of jgm_model_size(kind) doubles, row-major (device); the vertical-gradient kind's
 * background columns are zero (OMPMagneticGradImp.h:27-34), the magnetization kind's
 * columns are [Mx; My; Mz]. */
jgm_status jgm_sensitivity_rows(jgm_context *ctx, jgm_kind kind, size_t site0, size_t nsites,
                                double *rows);

/* Stored J in HBM, FP32, compact layout: rows = nsites * k with, per site,
 * gz (scalar, k = 1); Uxx, Uxy, Uxz, Uyy, Uyz, Uzz (tensor, k = 6); Bx, By, Bz per unit
 * susceptibility (magnetic, k = 3); gz then the six tensor values (scalar+tensor, k = 7);
 * Bz at the site minus Bz 1 m below (vertical gradient, k = 1); Bx, By, Bz
 * (magnetization, k = 3).  Columns = ncells (background layers are not stored; the
 * vertical-gradient kind's background gradient is zero, as in OMPMagneticGradImp), or
 * 3 ncells [Mx; My; Mz] for the magnetization kind; column-major (rows fastest).
 * Replaces the ublas matrix of FullSensitivityGravMagCalculator (:40-41, :116-130).
 * Every entry is evaluated with the FP64 closed form and rounded once to FP32, so
 * prec (JGM_FP64, or JGM_MIXED for the kinds that have it) does not change J. */
jgm_status jgm_storedj_build(jgm_context *ctx, jgm_kind kind, jgm_precision prec);
size_t jgm_storedj_rows(const jgm_context *ctx);
/* Y = J M for nrhs model vectors (M: cols x nrhs column-major with cols the stored J's
 * column count, ncells or 3 ncells for the magnetization kind; Y: rows x nrhs
 * column-major; device pointers), batching CalculateRawData (:137-152) over L-BFGS
 * trial models.  JGM_GEMM_FP32 multiplies FP32 operands; JGM_GEMM_BF16 rounds J and M
 * to BF16 (FP32 accumulation); JGM_GEMM_BF16_REFINED adds the products of the BF16
 * remainders (hi*hi + hi*lo + lo*hi), close to FP32 accuracy at BF16 matrix-core rate.
 * The first JGM_GEMM_BF16 product of an HBM-resident J makes a BF16 copy of J (2 bytes
 * per element of device memory, rows padded to a multiple of 256, kept until J is
 * rebuilt or released) that the J M products read instead of the FP32 J, with the same
 * results; when that memory cannot be allocated they read the FP32 J. */
jgm_status jgm_storedj_apply(jgm_context *ctx, jgm_gemm_precision prec, const float *M,
                             size_t nrhs, float *Y);
/* G = J^T R for nrhs residual vectors (R: rows x nrhs, G: cols x nrhs), the batched
 * form of CalculateRawLQDerivative (:214-221) without its factor 2. */
jgm_status jgm_storedj_apply_t(jgm_context *ctx, jgm_gemm_precision prec, const float *R,
                               size_t nrhs, float *G);
/* Cached forward over the stored J: the kind's data (transformed when a transform is
 * set) of model = n_params FP64 values (cells, then background layers), i.e.
 * FullSensitivityGravMagCalculator::CalculateCachedResult (:154-164) with the raw data
 * J m of CalculateRawData (:137-152); the background columns come from the cached
 * layer table.  J m runs on the matrix cores in `prec` (FP32 operands: model rounded
 * to FP32, totals of at most 256 cells accumulated in FP32 and combined in FP64).
 * Device pointers; the _host variants take host arrays. */
jgm_status jgm_storedj_forward(jgm_context *ctx, jgm_gemm_precision prec, const double *model,
                               double *out);
/* Cached gradient 2 [J B]^T D^T misfit (CachedLQDerivative :166-211 and
 * CalculateRawLQDerivative :214-221; D is the transform Jacobian at the model's data
 * through the stored J, needed only with a transform). */
jgm_status jgm_storedj_gradient(jgm_context *ctx, jgm_gemm_precision prec, const double *model,
                                const double *misfit, double *grad);
jgm_status jgm_storedj_forward_host(jgm_context *ctx, jgm_gemm_precision prec,
                                    const double *model, double *out);
jgm_status jgm_storedj_gradient_host(jgm_context *ctx, jgm_gemm_precision prec,
                                     const double *model, const double *misfit, double *grad);
/* Batched cached forward and gradient for several models (line-search trials, see
 * jgm_forward_batch): up to 64 models share one pass over J, because the products'
 * right-hand-side tile is 64 wide, so a batch of 64 costs what one model costs (the
 * FullSensitivityGravMagCalculator of the reference does one prec_prod per model,
 * FullSensitivityGravMagCalculator.h:151, :220).  Leading dimensions and the host forms
 * as for jgm_forward_batch / jgm_adjoint_batch; results equal the single-model calls
 * bit for bit. */
jgm_status jgm_storedj_forward_batch(jgm_context *ctx, jgm_gemm_precision prec, size_t nmodels,
                                     const double *models, size_t model_ld, double *out,
                                     size_t out_ld);
jgm_status jgm_storedj_gradient_batch(jgm_context *ctx, jgm_gemm_precision prec, size_t nmodels,
                                      const double *models, size_t model_ld,
                                      const double *misfits, size_t misfit_ld, double *grads,
                                      size_t grad_ld);
jgm_status jgm_storedj_forward_batch_host(jgm_context *ctx, jgm_gemm_precision prec,
                                          size_t nmodels, const double *models, double *out);
jgm_status jgm_storedj_gradient_batch_host(jgm_context *ctx, jgm_gemm_precision prec,
                                           size_t nmodels, const double *models,
                                           const double *misfits, double *grads);
jgm_status jgm_storedj_release(jgm_context *ctx);

/* Stored J in FP64, the precision of the reference's FullSensitivity matrix (a
 * ublas::matrix<double>, FullSensitivityGravMagCalculator.h:40-41, :116-130): same compact
 * layout as jgm_storedj_build, 8 bytes per element (7.5 GB for 7168 x 131072).  The cached
 * forward and gradient (jgm_storedj_forward / _gradient and their batch and host forms)
 * then take JGM_GEMM_FP64: model and residuals stay FP64 and J m, J^T r run on the FP64
 * matrix cores with FP64 accumulation, so cached results equal the direct FP64 path to
 * FP64 rounding (the reference's own caching tests ask for float epsilon).  The other
 * GEMM precisions need the FP32 J and return JGM_ERR_UNSUPPORTED on an FP64 J, and
 * JGM_GEMM_FP64 on an FP32 J.  Held in HBM only (no host-resident form). */
jgm_status jgm_storedj_build_fp64(jgm_context *ctx, jgm_kind kind);
/* Y = J M and G = J^T R on the FP64 J: M cols x nrhs, Y rows x nrhs, R rows x nrhs,
 * G cols x nrhs, column-major FP64 device arrays; the FP64 forms of jgm_storedj_apply /
 * _apply_t (CalculateRawData :137-152 and CalculateRawLQDerivative :214-221 without its
 * factor 2, batched over right-hand sides). */
jgm_status jgm_storedj_apply_f64(jgm_context *ctx, const double *M, size_t nrhs, double *Y);
jgm_status jgm_storedj_apply_t_f64(jgm_context *ctx, const double *R, size_t nrhs, double *G);

/* Stored J for problems whose J does not fit in device memory: the role of the
 * reference's DiskGravMagCalculator (jif3D/GravMag/DiskGravMagCalculator.h:100-200),
 * which writes J to a file and reads it back for every cached product.  Here J (same
 * FP32 compact layout) is built in column blocks of block_cells cells (rounded up to a
 * multiple of 256) and kept in page-locked host memory; every product afterwards
 * streams the blocks over PCIe into two HBM staging buffers, the copy of one block
 * overlapping the matrix-core work on the previous one.  All stored-J entry points
 * (apply, apply_t, forward, gradient and their batch forms) work unchanged; J^T R
 * results equal the HBM-resident ones bit for bit, J M results differ only by the
 * grouping of the FP64 sums over blocks.  Device memory used: 2 * block_cells * rows
 * * 4 bytes plus the operands.  The reference's cached Disk gradient omits the factor 2
 * of LQDerivative (DiskGravMagCalculator.h:164-199); this one keeps it.
 * jgm_storedj_host_resident returns 1 when the current stored J lives in host memory. */
jgm_status jgm_storedj_build_host(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                                  size_t block_cells);
int jgm_storedj_host_resident(const jgm_context *ctx);
/* Stored J larger than host memory too: as jgm_storedj_build_host, but J is written to a
 * file in `directory` (the reference's DiskGravMagCalculator writes its J to a file in a
 * temporary directory and reads it back for every cached product, DiskGravMagCalculator.h:
 * 76-85, :127-200).  Every product reads the blocks back into two page-locked buffers, the
 * read of block b + 1 overlapping the copy to HBM and the matrix-core work of block b;
 * results equal the host-resident J's bit for bit.  The file is removed when J is released,
 * rebuilt or the context destroyed.  Host memory used: 2 * block_cells * rows * 4 bytes.
 * jgm_storedj_file_backed returns 1 when the current stored J is in a file. */
jgm_status jgm_storedj_build_file(jgm_context *ctx, jgm_kind kind, jgm_precision prec,
                                  size_t block_cells, const char *directory);
int jgm_storedj_file_backed(const jgm_context *ctx);

/* Device memory on the context's device, for callers without their own HIP
 * allocator (the CUDA host wrappers did this inside PrepareData/FreeData,
 * jif3D/Gravity/gravcuda.cu:8-107).  Copies are ordered on the context's stream
 * and return when complete. */
jgm_status jgm_device_alloc(jgm_context *ctx, size_t bytes, void **ptr);
jgm_status jgm_device_free(jgm_context *ctx, void *ptr);
jgm_status jgm_copy_to_device(jgm_context *ctx, void *dst, const void *src, size_t bytes);
jgm_status jgm_copy_to_host(jgm_context *ctx, void *dst, const void *src, size_t bytes);
jgm_status jgm_copy_device(jgm_context *ctx, void *dst, const void *src, size_t bytes);

/* Kernel timing with HIP events recorded on the context's stream around every
 * kernel launch while enabled (replaces the wall-clock harness of
 * jif3D/Profiling/time_gravity_size.cpp:107-157).  jgm_profile_get reports the
 * summed time and launch count of entry `index` (0 .. jgm_profile_count-1);
 * the name pointer stays valid until the next reset. */
jgm_status jgm_profile_enable(jgm_context *ctx, int enable);
jgm_status jgm_profile_reset(jgm_context *ctx);
size_t jgm_profile_count(jgm_context *ctx);
jgm_status jgm_profile_get(jgm_context *ctx, size_t index, const char **name, double *total_ms,
                           size_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* JIF3D_GRAVMAG_H_ */
