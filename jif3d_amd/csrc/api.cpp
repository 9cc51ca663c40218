// C-ABI layer: argument checking, context state, staging of host arrays and the
// dispatch to the HIP kernels.  See include/jif3d_gravmag.h for the contract.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "context.hpp"
#include "jif3d_gravmag.h"

namespace {

thread_local std::string g_last_error;

jgm_status fail(jgm_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return st;
}

jgm_status hip_fail(hipError_t e, const char* where) {
  // a failed runtime call also leaves its code as the thread's last error; consume it so
  // the next launch check of an unrelated call does not report it again
  (void)hipGetLastError();
  if (e == hipErrorOutOfMemory)
    return fail(JGM_ERR_OUT_OF_MEMORY, "%s: out of device memory", where);
  return fail(JGM_ERR_DEVICE, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
}

// Makes the context's device current for the duration of a call.
struct DeviceGuard {
  int prev = -1;
  hipError_t err = hipSuccess;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) err = hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

bool kind_ok(int kind) {
  return kind == JGM_KIND_SCALAR || kind == JGM_KIND_TENSOR || kind == JGM_KIND_MAGNETIC ||
         kind == JGM_KIND_SCALAR_TENSOR || kind == JGM_KIND_MAGNETIC_GRADIENT ||
         kind == JGM_KIND_MAGNETIZATION;
}

jgm_status check_ready(const jgm_context* ctx, int kind, int prec) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (!kind_ok(kind)) return fail(JGM_ERR_INVALID_ARGUMENT, "unknown kind %d", kind);
  if (prec != JGM_FP64 && prec != JGM_MIXED)
    return fail(JGM_ERR_INVALID_ARGUMENT, "unknown precision %d", prec);
  if (!ctx->mesh_set) return fail(JGM_ERR_NOT_READY, "mesh not set (jgm_set_mesh)");
  if (!ctx->sites_set) return fail(JGM_ERR_NOT_READY, "sites not set (jgm_set_sites)");
  const int tf = ctx->transform;
  if (tf == JGM_TF_FTG_INVARIANT && kind != JGM_KIND_TENSOR)
    return fail(JGM_ERR_INVALID_ARGUMENT, "the FTG invariant transform needs the tensor kind");
  if ((tf == JGM_TF_TOTAL_FIELD || tf == JGM_TF_TOTAL_FIELD_ANOMALY) &&
      kind != JGM_KIND_MAGNETIC && kind != JGM_KIND_MAGNETIZATION)
    return fail(JGM_ERR_INVALID_ARGUMENT,
                "total-field transforms need the magnetic or magnetization kind");
  if (jgm::ncells(ctx) >= (size_t)1 << 31 || jgm::ncorners(ctx) >= ((size_t)1 << 40))
    return fail(JGM_ERR_UNSUPPORTED, "mesh too large");
  return JGM_OK;
}

jgm_status upload(jgm::DeviceBuffer& b, const double* h, size_t n, hipStream_t st,
                  const char* what) {
  hipError_t e = b.ensure(sizeof(double) * (n ? n : 1));
  if (e != hipSuccess) return hip_fail(e, what);
  if (n) {
    e = hipMemcpyAsync(b.ptr, h, sizeof(double) * n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, what);
  }
  return JGM_OK;
}

// After a mixed forward: FP64 at the sample sites against the mixed raw data; above the
// tolerance the whole forward is recomputed in FP64.
jgm_status mixed_forward_check(jgm_context* ctx, int kind, const double* model, double* out,
                               double* raw_out) {
  if (!jgm::mixed_check_applies(ctx, kind)) return JGM_OK;
  double err = 0.0;
  hipError_t e = jgm::launch_mixed_check_forward(ctx, kind, model, raw_out ? raw_out : out, &err);
  if (e != hipSuccess) return hip_fail(e, "forward (mixed check)");
  ctx->check_last_fwd = err;
  if (err > ctx->check_tol) {
    ++ctx->check_fallbacks;
    if ((e = jgm::launch_forward_fp64(ctx, kind, model, out, raw_out)) != hipSuccess)
      return hip_fail(e, "forward (FP64 after the mixed check)");
  }
  return JGM_OK;
}

// After a mixed adjoint (raw misfit rm): FP64 at the sample cells against the mixed
// gradient; above the tolerance the gradient is recomputed in FP64.
jgm_status mixed_adjoint_check(jgm_context* ctx, int kind, const double* rm, double* grad) {
  if (!jgm::mixed_check_applies(ctx, kind)) return JGM_OK;
  double err = 0.0;
  hipError_t e = jgm::launch_mixed_check_adjoint(ctx, kind, rm, grad, &err);
  if (e != hipSuccess) return hip_fail(e, "adjoint (mixed check)");
  ctx->check_last_adj = err;
  if (err > ctx->check_tol) {
    ++ctx->check_fallbacks;
    if ((e = jgm::launch_adjoint_fp64(ctx, kind, nullptr, rm, grad)) != hipSuccess)
      return hip_fail(e, "adjoint (FP64 after the mixed check)");
  }
  return JGM_OK;
}

jgm_status ensure_background(jgm_context* ctx) {
  if (ctx->bg_valid) return JGM_OK;
  hipError_t e = jgm::launch_background_table(ctx);
  if (e != hipSuccess) return hip_fail(e, "background table");
  return JGM_OK;
}

size_t data_per_site(const jgm_context* ctx, int kind) {
  if (!kind_ok(kind)) return 0;
  return ctx->transform == JGM_TF_NONE ? jgm::raw_per_site(kind) : 1;
}

// Host-memory form of a device entry point: copies the inputs into temporary device
// buffers, runs `call` on their device pointers, copies the outputs back and waits.
struct HostArg {
  const double* in;  // host input (or nullptr)
  double* out;       // host output (or nullptr)
  size_t n;          // doubles
};
template <size_t NA, class F>
jgm_status run_staged(jgm_context* ctx, const char* where, const HostArg (&args)[NA], F call) {
  jgm::DeviceBuffer buf[NA];
  const double* dev_in[NA];
  double* dev_out[NA];
  hipError_t e = hipSuccess;
  for (size_t a = 0; a < NA && e == hipSuccess; ++a) {
    dev_in[a] = nullptr;
    dev_out[a] = nullptr;
    if (!args[a].in && !args[a].out) continue;
    e = buf[a].ensure(sizeof(double) * (args[a].n ? args[a].n : 1));
    if (e == hipSuccess && args[a].in && args[a].n)
      e = hipMemcpyAsync(buf[a].ptr, args[a].in, sizeof(double) * args[a].n,
                         hipMemcpyHostToDevice, ctx->stream);
    dev_in[a] = static_cast<const double*>(buf[a].ptr);
    dev_out[a] = static_cast<double*>(buf[a].ptr);
  }
  jgm_status st = JGM_OK;
  if (e == hipSuccess) {
    st = call(dev_in, dev_out);
    for (size_t a = 0; a < NA && st == JGM_OK && e == hipSuccess; ++a)
      if (args[a].out && args[a].n)
        e = hipMemcpyAsync(args[a].out, buf[a].ptr, sizeof(double) * args[a].n,
                           hipMemcpyDeviceToHost, ctx->stream);
  }
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  for (size_t a = 0; a < NA; ++a) buf[a].release();
  if (st != JGM_OK) return st;
  if (e != hipSuccess) return hip_fail(e, where);
  if (e2 != hipSuccess) return hip_fail(e2, where);
  return JGM_OK;
}

}  // namespace

extern "C" {

const char* jgm_last_error(void) { return g_last_error.c_str(); }

int jgm_api_version(void) { return JGM_API_VERSION; }

size_t jgm_raw_per_site(jgm_kind kind) { return jgm::raw_per_site(kind); }

jgm_status jgm_create(int device, void* stream, jgm_context** out) {
  if (!out) return fail(JGM_ERR_INVALID_ARGUMENT, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (device < 0 || device >= ndev)
    return fail(JGM_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, ndev);
  DeviceGuard g(device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  jgm_context* ctx = new (std::nothrow) jgm_context();
  if (!ctx) return fail(JGM_ERR_OUT_OF_MEMORY, "host allocation failed");
  ctx->device = device;
  ctx->stream = static_cast<hipStream_t>(stream);
  if (const char* v = std::getenv("JIF3D_AMD_BF16_COPY")) ctx->bf16_copy_enabled = v[0] != '0';
  if (const char* v = std::getenv("JIF3D_AMD_FAR_FORWARD_WAVES"))
    ctx->far_fwd_waves = std::max(1L, std::atol(v));
  if (const char* v = std::getenv("JIF3D_AMD_FAR_ADJOINT_WAVES"))
    ctx->far_adj_waves = std::max(1L, std::atol(v));
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    ctx->cu_count = cus;
  // the device address of the allocation-free forward's scratch array, once per device
  jgm::noalloc_scratch_init(device);
  *out = ctx;
  return JGM_OK;
}

jgm_status jgm_destroy(jgm_context* ctx) {
  if (!ctx) return JGM_OK;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  jgm::DeviceBuffer* bufs[] = {&ctx->d_Xc, &ctx->d_Yc, &ctx->d_Zc, &ctx->d_xlo, &ctx->d_ylo,
                               &ctx->d_zlo, &ctx->d_xsz, &ctx->d_ysz, &ctx->d_zsz, &ctx->d_sx,
                               &ctx->d_sy, &ctx->d_sz, &ctx->bgT, &ctx->near.box,
                               &ctx->weights, &ctx->partial, &ctx->raw,
                               &ctx->coef, &ctx->hpart, &ctx->hsum, &ctx->model_stage,
                               &ctx->out_stage, &ctx->misfit_stage, &ctx->grad_stage,
                               &ctx->farpart, &ctx->nearpart, &ctx->coef32, &ctx->storedJ, &ctx->storedJ_vec,
                               &ctx->storedJ_bf16, &ctx->storedJ_bf16c, &ctx->storedJ_rhs16lo,
                               &ctx->near.col_ptr, &ctx->near.col_sites, &ctx->near.col_ids,
                               &ctx->near.col_kmax, &ctx->near.xcen, &ctx->near.ycen,
                               &ctx->near.zcen, &ctx->near.adj_table, &ctx->near.corner_off,
                               &ctx->near.cell_off, &ctx->near_h, &ctx->near_v, &ctx->vsites, &ctx->coef_mono,
                               &ctx->near.site_order, &ctx->near.xoff,
                               &ctx->near.yoff, &ctx->near.zoff, &ctx->near.kmap,
                               &ctx->near.listed, &ctx->acc64, &ctx->raw_cache_model, &ctx->raw_cache_raw, &ctx->raw_cache_flag,
                               &ctx->partsum, &ctx->chunk_scratch, &ctx->mag_part};
  for (auto* b : bufs) b->release();
  jgm::release_slab_mesh(ctx);
  jgm::release_host_storedj(ctx);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  for (hipEvent_t ev : {ctx->ev_copied[0], ctx->ev_copied[1], ctx->ev_used[0], ctx->ev_used[1],
                        ctx->ev_start})
    if (ev) (void)hipEventDestroy(ev);
  jgm::prof_clear(ctx);
  delete ctx;
  return JGM_OK;
}

jgm_status jgm_set_stream(jgm_context* ctx, void* stream) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  ctx->stream = static_cast<hipStream_t>(stream);
  return JGM_OK;
}

jgm_status jgm_synchronize(jgm_context* ctx) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  DeviceGuard g(ctx->device);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return JGM_OK;
}

jgm_status jgm_set_mesh(jgm_context* ctx, size_t nx, size_t ny, size_t nz, const double* xcoords,
                        const double* ycoords, const double* zcoords, const double* xsizes,
                        const double* ysizes, const double* zsizes) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  const size_t n[3] = {nx, ny, nz};
  const double* c[3] = {xcoords, ycoords, zcoords};
  const double* s[3] = {xsizes, ysizes, zsizes};
  for (int a = 0; a < 3; ++a) {
    if (n[a] > 0 && (!c[a] || !s[a]))
      return fail(JGM_ERR_INVALID_ARGUMENT, "axis %d: null coordinate or size array", a);
    if (n[a] > (1u << 20)) return fail(JGM_ERR_UNSUPPORTED, "axis %d: too many cells", a);
    for (size_t i = 0; i < n[a]; ++i)
      if (!std::isfinite(s[a][i]) || !std::isfinite(c[a][i]))
        return fail(JGM_ERR_INVALID_ARGUMENT, "axis %d: non-finite coordinate or size", a);
  }
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  ctx->nx = nx;
  ctx->ny = ny;
  ctx->nz = nz;
  std::vector<double>* hc[3] = {&ctx->xc, &ctx->yc, &ctx->zc};
  std::vector<double>* hs[3] = {&ctx->xs, &ctx->ys, &ctx->zs};
  double width[3];
  for (int a = 0; a < 3; ++a) {
    hc[a]->assign(n[a] + 1, 0.0);
    if (n[a] > 0)
      hs[a]->assign(s[a], s[a] + n[a]);
    else
      hs[a]->clear();
    // corner a of the mesh: cell i spans [coord[i], coord[i] + size[i]]; the last
    // corner is the far face of the last cell as CalcGravBoxTerm forms it.
    if (n[a] > 0) {
      for (size_t i = 0; i < n[a]; ++i) (*hc[a])[i] = c[a][i];
      (*hc[a])[n[a]] = c[a][n[a] - 1] + s[a][n[a] - 1];
    } else {
      (*hc[a])[0] = c[a] ? c[a][0] : 0.0;
    }
    double w = 0.0;  // std::accumulate(sizes, 0.0), ThreeDGravMagImplementation.h:132-137
    for (size_t i = 0; i < n[a]; ++i) w += s[a][i];
    width[a] = w;
  }
  ctx->xwidth = width[0];
  ctx->ywidth = width[1];
  ctx->zwidth = width[2];
  jgm_status st;
  if ((st = upload(ctx->d_Xc, ctx->xc.data(), nx + 1, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_Yc, ctx->yc.data(), ny + 1, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_Zc, ctx->zc.data(), nz + 1, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_xlo, ctx->xc.data(), nx, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_ylo, ctx->yc.data(), ny, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_zlo, ctx->zc.data(), nz, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_xsz, ctx->xs.data(), nx, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_ysz, ctx->ys.data(), ny, ctx->stream, "mesh")) != JGM_OK) return st;
  if ((st = upload(ctx->d_zsz, ctx->zs.data(), nz, ctx->stream, "mesh")) != JGM_OK) return st;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "mesh upload");
  ++ctx->config_gen;
  ++ctx->geom_gen;
  ctx->mesh_set = true;
  ctx->bg_valid = false;
  ctx->near.valid = false;
  ctx->storedJ_kind = 0;
  return JGM_OK;
}

jgm_status jgm_set_background(jgm_context* ctx, size_t nlayers, const double* thicknesses) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (nlayers > 0 && !thicknesses) return fail(JGM_ERR_INVALID_ARGUMENT, "null thicknesses");
  if (nlayers > 4096) return fail(JGM_ERR_UNSUPPORTED, "too many background layers");
  ++ctx->config_gen;
  ctx->bgthick.assign(thicknesses, thicknesses + nlayers);
  ctx->bg_valid = false;
  ctx->storedJ_kind = 0;
  return JGM_OK;
}

jgm_status jgm_set_sites(jgm_context* ctx, size_t nsites, const double* x, const double* y,
                         const double* z) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (nsites > 0 && (!x || !y || !z)) return fail(JGM_ERR_INVALID_ARGUMENT, "null site array");
  if (nsites >= (1u << 30)) return fail(JGM_ERR_UNSUPPORTED, "too many sites");
  for (size_t i = 0; i < nsites; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]))
      return fail(JGM_ERR_INVALID_ARGUMENT, "site %zu has a non-finite coordinate", i);
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  ctx->S = nsites;
  ctx->hsx.assign(x, x + nsites);
  ctx->hsy.assign(y, y + nsites);
  ctx->hsz.assign(z, z + nsites);
  jgm_status st;
  if ((st = upload(ctx->d_sx, x, nsites, ctx->stream, "sites")) != JGM_OK) return st;
  if ((st = upload(ctx->d_sy, y, nsites, ctx->stream, "sites")) != JGM_OK) return st;
  if ((st = upload(ctx->d_sz, z, nsites, ctx->stream, "sites")) != JGM_OK) return st;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "site upload");
  ++ctx->config_gen;
  ++ctx->geom_gen;
  ctx->sites_set = true;
  ctx->bg_valid = false;
  ctx->near.valid = false;
  ctx->storedJ_kind = 0;
  return JGM_OK;
}

jgm_status jgm_set_field(jgm_context* ctx, double inclination, double declination,
                         double fieldstrength) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (!std::isfinite(inclination) || !std::isfinite(declination) || !std::isfinite(fieldstrength))
    return fail(JGM_ERR_INVALID_ARGUMENT, "non-finite field parameter");
  const bool changed = inclination != ctx->inclination || declination != ctx->declination ||
                       fieldstrength != ctx->fieldstrength;
  ++ctx->config_gen;
  ctx->inclination = inclination;
  ctx->declination = declination;
  ctx->fieldstrength = fieldstrength;
  // the stored J of the susceptibility and vertical-gradient kinds holds the field
  // direction (and the gradient kind its strength): it no longer applies
  if (changed &&
      (ctx->storedJ_kind == JGM_KIND_MAGNETIC || ctx->storedJ_kind == JGM_KIND_MAGNETIC_GRADIENT))
    ctx->storedJ_kind = 0;
  return JGM_OK;
}

jgm_status jgm_set_transform(jgm_context* ctx, jgm_transform tf, double inclination,
                             double declination) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (tf != JGM_TF_NONE && tf != JGM_TF_FTG_INVARIANT && tf != JGM_TF_TOTAL_FIELD &&
      tf != JGM_TF_TOTAL_FIELD_ANOMALY)
    return fail(JGM_ERR_INVALID_ARGUMENT, "unknown transform %d", (int)tf);
  ++ctx->config_gen;
  ctx->transform = tf;
  ctx->tf_inc = inclination;
  ctx->tf_dec = declination;
  return JGM_OK;
}

jgm_status jgm_set_mixed_threshold(jgm_context* ctx, double q0) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (!(q0 >= 2.0 && q0 <= 1e6)) return fail(JGM_ERR_INVALID_ARGUMENT, "q0 must be in [2, 1e6]");
  ++ctx->config_gen;
  ctx->mixed_q0 = q0;
  ctx->near.valid = false;
  return JGM_OK;
}

jgm_status jgm_set_mixed_far_radius(jgm_context* ctx, double q1) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (!(q1 == 0.0 || (q1 >= 2.0 && q1 <= 1e6)))
    return fail(JGM_ERR_INVALID_ARGUMENT, "q1 must be 0 or in [2, 1e6]");
  ++ctx->config_gen;
  ctx->mixed_q1 = q1;
  return JGM_OK;
}

jgm_status jgm_set_mixed_check(jgm_context* ctx, size_t nsites, size_t ncells,
                               double tolerance) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (!(tolerance > 0.0) || !std::isfinite(tolerance))
    return fail(JGM_ERR_INVALID_ARGUMENT, "the check tolerance must be positive");
  if (nsites > 4096 || ncells > 65536)
    return fail(JGM_ERR_UNSUPPORTED, "at most 4096 sample sites and 65536 sample cells");
  ctx->check_sites = nsites;
  ctx->check_cells = ncells;
  ctx->check_tol = tolerance;
  ctx->check_gen = 0;  // new samples
  return JGM_OK;
}

jgm_status jgm_mixed_check_report(const jgm_context* ctx, double* forward_error,
                                  double* adjoint_error, size_t* fallbacks) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (forward_error) *forward_error = ctx->check_last_fwd;
  if (adjoint_error) *adjoint_error = ctx->check_last_adj;
  if (fallbacks) *fallbacks = ctx->check_fallbacks;
  return JGM_OK;
}

size_t jgm_data_per_site(const jgm_context* ctx, jgm_kind kind) {
  return ctx ? data_per_site(ctx, kind) : 0;
}

size_t jgm_n_params(const jgm_context* ctx) { return ctx ? jgm::nparams(ctx) : 0; }

size_t jgm_model_size(const jgm_context* ctx, jgm_kind kind) {
  return ctx && kind_ok(kind) ? jgm::nparams(ctx, kind) : 0;
}

jgm_status jgm_forward(jgm_context* ctx, jgm_kind kind, jgm_precision prec, const double* model,
                       double* out) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (!model || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null model or output pointer");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if (ctx->S == 0) return JGM_OK;
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e;
  // with a transform, keep the raw data and the model for an adjoint at the same model
  double* raw_out = nullptr;
  const size_t nmodel = jgm::nparams(ctx, kind), nraw = ctx->S * jgm::raw_per_site(kind);
  ctx->raw_cache.valid = false;
  if (ctx->transform != JGM_TF_NONE) {
    if ((e = ctx->raw_cache_raw.ensure(sizeof(double) * nraw)) != hipSuccess ||
        (e = ctx->raw_cache_model.ensure(sizeof(double) * (nmodel ? nmodel : 1))) != hipSuccess)
      return hip_fail(e, "forward");
    raw_out = ctx->raw_cache_raw.as<double>();
  }
  if (prec == JGM_MIXED)
    e = jgm::launch_forward_mixed(ctx, kind, model, out, raw_out);
  else
    e = jgm::launch_forward_fp64(ctx, kind, model, out, raw_out);
  if (e != hipSuccess) return hip_fail(e, "forward");
  if (prec == JGM_MIXED && (st = mixed_forward_check(ctx, kind, model, out, raw_out)) != JGM_OK)
    return st;
  if (raw_out) {
    if (nmodel && (e = hipMemcpyAsync(ctx->raw_cache_model.ptr, model, sizeof(double) * nmodel,
                                      hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess)
      return hip_fail(e, "forward");
    ctx->raw_cache = {true, (int)kind, (int)prec, ctx->config_gen, nmodel};
  }
  return JGM_OK;
}

jgm_status jgm_forward_noalloc(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                               const double* model, double* out) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (prec != JGM_FP64)
    return fail(JGM_ERR_UNSUPPORTED, "jgm_forward_noalloc runs in JGM_FP64 only");
  if (kind == JGM_KIND_MAGNETIC_GRADIENT || kind == JGM_KIND_MAGNETIZATION)
    return fail(JGM_ERR_UNSUPPORTED,
                "jgm_forward_noalloc: the scalar, tensor, magnetic and scalar+tensor kinds only");
  if (!model || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null model or output pointer");
  if (ctx->S == 0) return JGM_OK;
  const size_t need = jgm::noalloc_forward_need(ctx, kind);
  if (need > (size_t)JGM_NOALLOC_SCRATCH_BYTES)
    return fail(JGM_ERR_UNSUPPORTED,
                "jgm_forward_noalloc: this mesh and survey need %zu bytes of scratch "
                "(corner weights, background table, one chunk of partial sums); the "
                "allocation-free entry point has %zu",
                need, (size_t)JGM_NOALLOC_SCRATCH_BYTES);
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  hipError_t e = jgm::launch_forward_fp64_noalloc(ctx, kind, model, out);
  if (e == hipErrorOutOfMemory)
    return fail(JGM_ERR_UNSUPPORTED, "jgm_forward_noalloc: scratch of %zu bytes too small",
                (size_t)JGM_NOALLOC_SCRATCH_BYTES);
  if (e != hipSuccess) return hip_fail(e, "jgm_forward_noalloc");
  return JGM_OK;
}

jgm_status jgm_adjoint(jgm_context* ctx, jgm_kind kind, jgm_precision prec, const double* model,
                       const double* misfit, double* grad) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (!misfit || !grad) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  if (ctx->transform != JGM_TF_NONE && !model)
    return fail(JGM_ERR_INVALID_ARGUMENT, "a transform needs the model for its Jacobian");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  const size_t S = ctx->S;
  const size_t kraw = jgm::raw_per_site(kind);
  hipError_t e;
  if (S == 0) {
    e = jgm::launch_fill_zero(ctx, grad, jgm::nparams(ctx, kind));
    if (e != hipSuccess) return hip_fail(e, "adjoint");
    return JGM_OK;
  }
  const double* rm = misfit;
  if (ctx->transform != JGM_TF_NONE) {
    // raw data at the model for the transform Jacobian (LQDerivative recomputes it
    // per site, ThreeDGravMagImplementation.h:209-214): taken from the last forward when
    // that ran with the same kind, precision and settings on a bitwise equal model
    if ((e = ctx->misfit_stage.ensure(sizeof(double) * S * kraw)) != hipSuccess)
      return hip_fail(e, "adjoint");
    const double* raw = nullptr;
    const jgm_context::RawCache& rc = ctx->raw_cache;
    if (rc.valid && rc.kind == (int)kind && rc.prec == (int)prec && rc.gen == ctx->config_gen &&
        rc.nmodel == jgm::nparams(ctx, kind)) {
      int differ = 1;
      if ((e = ctx->raw_cache_flag.ensure(sizeof(int))) != hipSuccess ||
          (e = jgm::launch_models_differ(ctx, model, ctx->raw_cache_model.as<const double>(),
                                         rc.nmodel, ctx->raw_cache_flag.as<int>())) != hipSuccess ||
          (e = hipMemcpyAsync(&differ, ctx->raw_cache_flag.ptr, sizeof(int),
                              hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
          (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return hip_fail(e, "adjoint (model comparison)");
      if (!differ) raw = ctx->raw_cache_raw.as<const double>();
    }
    if (!raw) {
      if ((e = ctx->raw.ensure(sizeof(double) * S * kraw)) != hipSuccess)
        return hip_fail(e, "adjoint");
      if ((e = ctx->out_stage.ensure(sizeof(double) * S)) != hipSuccess)
        return hip_fail(e, "adjoint");
      if (prec == JGM_MIXED)
        e = jgm::launch_forward_mixed(ctx, kind, model, ctx->out_stage.as<double>(),
                                      ctx->raw.as<double>());
      else
        e = jgm::launch_forward_fp64(ctx, kind, model, ctx->out_stage.as<double>(),
                                     ctx->raw.as<double>());
      if (e != hipSuccess) return hip_fail(e, "adjoint (raw forward)");
      raw = ctx->raw.as<const double>();
    }
    e = jgm::launch_misfit_to_raw(ctx, kind, misfit, raw, ctx->misfit_stage.as<double>());
    if (e != hipSuccess) return hip_fail(e, "adjoint (transform)");
    rm = ctx->misfit_stage.as<const double>();
  }
  if (prec == JGM_MIXED)
    e = jgm::launch_adjoint_mixed(ctx, kind, model, rm, grad);
  else
    e = jgm::launch_adjoint_fp64(ctx, kind, model, rm, grad);
  if (e != hipSuccess) return hip_fail(e, "adjoint");
  if (prec == JGM_MIXED) return mixed_adjoint_check(ctx, kind, rm, grad);
  return JGM_OK;
}

jgm_status jgm_forward_host(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                            const double* model, double* out) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (!model || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null model or output pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, kind);
  const size_t nout = ctx->S * data_per_site(ctx, kind);
  hipError_t e;
  if ((e = ctx->model_stage.ensure(sizeof(double) * (np ? np : 1))) != hipSuccess)
    return hip_fail(e, "forward_host");
  if ((e = ctx->grad_stage.ensure(sizeof(double) * (nout ? nout : 1))) != hipSuccess)
    return hip_fail(e, "forward_host");
  if (np && (e = hipMemcpyAsync(ctx->model_stage.ptr, model, sizeof(double) * np,
                                hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return hip_fail(e, "forward_host");
  st = jgm_forward(ctx, kind, prec, ctx->model_stage.as<double>(), ctx->grad_stage.as<double>());
  if (st != JGM_OK) return st;
  if (nout && (e = hipMemcpyAsync(out, ctx->grad_stage.ptr, sizeof(double) * nout,
                                  hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
    return hip_fail(e, "forward_host");
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(e, "forward_host");
  return JGM_OK;
}

jgm_status jgm_adjoint_host(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                            const double* model, const double* misfit, double* grad) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (!misfit || !grad) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, kind);
  const size_t nmis = ctx->S * data_per_site(ctx, kind);
  jgm::DeviceBuffer dmis;  // separate from the internal misfit_stage used by transforms
  hipError_t e;
  if ((e = ctx->model_stage.ensure(sizeof(double) * (np ? np : 1))) != hipSuccess)
    return hip_fail(e, "adjoint_host");
  if ((e = ctx->grad_stage.ensure(sizeof(double) * (np ? np : 1))) != hipSuccess)
    return hip_fail(e, "adjoint_host");
  if ((e = dmis.ensure(sizeof(double) * (nmis ? nmis : 1))) != hipSuccess)
    return hip_fail(e, "adjoint_host");
  if (model && np &&
      (e = hipMemcpyAsync(ctx->model_stage.ptr, model, sizeof(double) * np, hipMemcpyHostToDevice,
                          ctx->stream)) != hipSuccess) {
    dmis.release();
    return hip_fail(e, "adjoint_host");
  }
  if (nmis && (e = hipMemcpyAsync(dmis.ptr, misfit, sizeof(double) * nmis, hipMemcpyHostToDevice,
                                  ctx->stream)) != hipSuccess) {
    dmis.release();
    return hip_fail(e, "adjoint_host");
  }
  st = jgm_adjoint(ctx, kind, prec, model ? ctx->model_stage.as<double>() : nullptr,
                   dmis.as<double>(), ctx->grad_stage.as<double>());
  if (st == JGM_OK && np)
    e = hipMemcpyAsync(grad, ctx->grad_stage.ptr, sizeof(double) * np, hipMemcpyDeviceToHost,
                       ctx->stream);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  dmis.release();
  if (st != JGM_OK) return st;
  if (e != hipSuccess) return hip_fail(e, "adjoint_host");
  if (e2 != hipSuccess) return hip_fail(e2, "adjoint_host");
  return JGM_OK;
}

jgm_status jgm_forward_batch(jgm_context* ctx, jgm_kind kind, jgm_precision prec, size_t nmodels,
                             const double* models, size_t model_ld, double* out, size_t out_ld) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!models || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null models or output pointer");
  const size_t np = jgm::nparams(ctx, kind), nout = ctx->S * data_per_site(ctx, kind);
  if (nmodels > 1 && (model_ld < np || out_ld < nout))
    return fail(JGM_ERR_INVALID_ARGUMENT,
                "leading dimensions too small (model_ld %zu < %zu or out_ld %zu < %zu)", model_ld,
                np, out_ld, nout);
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if (ctx->S == 0) return JGM_OK;
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e = hipSuccess;
  if (prec == JGM_MIXED) {
    // with a transform the check compares raw data, kept here per model
    const bool check = jgm::mixed_check_applies(ctx, kind) && ctx->transform != JGM_TF_NONE;
    const size_t nraw = ctx->S * jgm::raw_per_site(kind);
    double* raw = nullptr;
    if (check) {
      if ((e = ctx->raw.ensure(sizeof(double) * nmodels * nraw)) != hipSuccess)
        return hip_fail(e, "forward_batch");
      raw = ctx->raw.as<double>();
    }
    e = jgm::launch_forward_mixed_batch(ctx, kind, nmodels, models, model_ld, out, out_ld, raw,
                                        nraw);
    if (e != hipSuccess) return hip_fail(e, "forward_batch");
    for (size_t b = 0; b < nmodels; ++b)
      if ((st = mixed_forward_check(ctx, kind, models + b * model_ld, out + b * out_ld,
                                    raw ? raw + b * nraw : nullptr)) != JGM_OK)
        return st;
    return JGM_OK;
  }
  e = jgm::launch_forward_fp64_batch(ctx, kind, nmodels, models, model_ld, out, out_ld, nullptr,
                                     0);
  if (e != hipSuccess) return hip_fail(e, "forward_batch");
  return JGM_OK;
}

jgm_status jgm_adjoint_batch(jgm_context* ctx, jgm_kind kind, jgm_precision prec, size_t nmodels,
                             const double* models, size_t model_ld, const double* misfits,
                             size_t misfit_ld, double* grads, size_t grad_ld) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!misfits || !grads) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  if (ctx->transform != JGM_TF_NONE && !models)
    return fail(JGM_ERR_INVALID_ARGUMENT, "a transform needs the models for its Jacobian");
  const size_t S = ctx->S, kraw = jgm::raw_per_site(kind);
  const size_t np = jgm::nparams(ctx, kind), nmis = S * data_per_site(ctx, kind);
  if (nmodels > 1 && (misfit_ld < nmis || grad_ld < np ||
                      (ctx->transform != JGM_TF_NONE && model_ld < np)))
    return fail(JGM_ERR_INVALID_ARGUMENT, "leading dimensions too small");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e = hipSuccess;
  if (S == 0) {
    for (size_t b = 0; b < nmodels && e == hipSuccess; ++b)
      e = jgm::launch_fill_zero(ctx, grads + b * grad_ld, np);
    if (e != hipSuccess) return hip_fail(e, "adjoint_batch");
    return JGM_OK;
  }
  const double* rm = misfits;
  size_t rm_ld = misfit_ld;
  if (ctx->transform != JGM_TF_NONE) {
    // raw data of every model for the transform Jacobians, then D^T r per model
    const size_t nraw = S * kraw;
    if ((e = ctx->raw.ensure(sizeof(double) * nmodels * nraw)) != hipSuccess ||
        (e = ctx->out_stage.ensure(sizeof(double) * nmodels * S)) != hipSuccess ||
        (e = ctx->misfit_stage.ensure(sizeof(double) * nmodels * nraw)) != hipSuccess)
      return hip_fail(e, "adjoint_batch");
    double* raw = ctx->raw.as<double>();
    double* outs = ctx->out_stage.as<double>();
    if (prec == JGM_MIXED) {
      e = jgm::launch_forward_mixed_batch(ctx, kind, nmodels, models, model_ld, outs, S, raw,
                                          nraw);
    } else {
      e = jgm::launch_forward_fp64_batch(ctx, kind, nmodels, models, model_ld, outs, S, raw,
                                         nraw);
    }
    if (e != hipSuccess) return hip_fail(e, "adjoint_batch (raw forward)");
    for (size_t b = 0; b < nmodels && e == hipSuccess; ++b)
      e = jgm::launch_misfit_to_raw(ctx, kind, misfits + b * misfit_ld, raw + b * nraw,
                                    ctx->misfit_stage.as<double>() + b * nraw);
    if (e != hipSuccess) return hip_fail(e, "adjoint_batch (transform)");
    rm = ctx->misfit_stage.as<const double>();
    rm_ld = nraw;
  }
  if (prec == JGM_MIXED) {
    e = jgm::launch_adjoint_mixed_batch(ctx, kind, nmodels, rm, rm_ld, grads, grad_ld);
    if (e != hipSuccess) return hip_fail(e, "adjoint_batch");
    for (size_t b = 0; b < nmodels; ++b)
      if ((st = mixed_adjoint_check(ctx, kind, rm + b * rm_ld, grads + b * grad_ld)) != JGM_OK)
        return st;
    return JGM_OK;
  }
  e = jgm::launch_adjoint_fp64_batch(ctx, kind, nmodels, rm, rm_ld, grads, grad_ld);
  if (e != hipSuccess) return hip_fail(e, "adjoint_batch");
  return JGM_OK;
}

jgm_status jgm_forward_batch_host(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                                  size_t nmodels, const double* models, double* out) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!models || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null models or output pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, kind), nout = ctx->S * data_per_site(ctx, kind);
  const HostArg args[2] = {{models, nullptr, nmodels * np}, {nullptr, out, nmodels * nout}};
  return run_staged(ctx, "forward_batch_host", args, [&](const double* const* in, double* const* o) {
    return jgm_forward_batch(ctx, kind, prec, nmodels, in[0], np, o[1], nout);
  });
}

jgm_status jgm_adjoint_batch_host(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                                  size_t nmodels, const double* models, const double* misfits,
                                  double* grads) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!misfits || !grads) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, kind), nmis = ctx->S * data_per_site(ctx, kind);
  const HostArg args[3] = {{models, nullptr, models ? nmodels * np : 0},
                           {misfits, nullptr, nmodels * nmis},
                           {nullptr, grads, nmodels * np}};
  return run_staged(ctx, "adjoint_batch_host", args, [&](const double* const* in, double* const* o) {
    return jgm_adjoint_batch(ctx, kind, prec, nmodels, models ? in[0] : nullptr, np, in[1], nmis,
                             o[2], np);
  });
}

jgm_status jgm_sensitivity_rows(jgm_context* ctx, jgm_kind kind, size_t site0, size_t nsites,
                                double* rows) {
  jgm_status st = check_ready(ctx, kind, JGM_FP64);
  if (st != JGM_OK) return st;
  if (!rows && nsites) return fail(JGM_ERR_INVALID_ARGUMENT, "null rows pointer");
  if (site0 > ctx->S || nsites > ctx->S - site0)
    return fail(JGM_ERR_INVALID_ARGUMENT, "site range [%zu, %zu) outside [0, %zu)", site0,
                site0 + nsites, ctx->S);
  if (nsites > 65535) return fail(JGM_ERR_UNSUPPORTED, "at most 65535 sites per call");
  if (nsites == 0) return JGM_OK;
  DeviceGuard g(ctx->device);
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e = jgm::launch_sensitivity_rows(ctx, kind, site0, nsites, rows);
  if (e != hipSuccess) return hip_fail(e, "sensitivity rows");
  return JGM_OK;
}

jgm_status jgm_storedj_build(jgm_context* ctx, jgm_kind kind, jgm_precision prec) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::launch_storedj_build(ctx, kind, prec);
  if (e == hipErrorNotSupported) return fail(JGM_ERR_UNSUPPORTED, "stored J not available");
  if (e != hipSuccess) return hip_fail(e, "stored J build");
  return JGM_OK;
}

jgm_status jgm_storedj_build_fp64(jgm_context* ctx, jgm_kind kind) {
  jgm_status st = check_ready(ctx, kind, JGM_FP64);
  if (st != JGM_OK) return st;
  if (ctx->S == 0 || jgm::ncells(ctx) == 0)
    return fail(JGM_ERR_INVALID_ARGUMENT, "no sites or no cells");
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::launch_storedj_build_f64(ctx, kind);
  if (e != hipSuccess) return hip_fail(e, "FP64 stored J build");
  return JGM_OK;
}

jgm_status jgm_storedj_build_host(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                                  size_t block_cells) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (block_cells == 0) return fail(JGM_ERR_INVALID_ARGUMENT, "block_cells must be positive");
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::launch_storedj_build_resident(ctx, kind, prec, block_cells);
  if (e != hipSuccess) return hip_fail(e, "host-resident stored J build");
  return JGM_OK;
}

int jgm_storedj_host_resident(const jgm_context* ctx) {
  return ctx && ctx->storedJ_kind && ctx->storedJ_host ? 1 : 0;
}

jgm_status jgm_storedj_build_file(jgm_context* ctx, jgm_kind kind, jgm_precision prec,
                                  size_t block_cells, const char* directory) {
  jgm_status st = check_ready(ctx, kind, prec);
  if (st != JGM_OK) return st;
  if (block_cells == 0) return fail(JGM_ERR_INVALID_ARGUMENT, "block_cells must be positive");
  if (!directory || !*directory) return fail(JGM_ERR_INVALID_ARGUMENT, "no directory");
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::launch_storedj_build_file(ctx, kind, block_cells, directory);
  if (e == hipErrorFileNotFound)
    return fail(JGM_ERR_DEVICE, "stored J file in %s: %s", directory, std::strerror(errno));
  if (e != hipSuccess) return hip_fail(e, "file-backed stored J build");
  return JGM_OK;
}

int jgm_storedj_file_backed(const jgm_context* ctx) {
  return ctx && ctx->storedJ_kind && ctx->storedJ_fd >= 0 ? 1 : 0;
}

size_t jgm_storedj_rows(const jgm_context* ctx) {
  return ctx && ctx->storedJ_kind ? ctx->storedJ_rows : 0;
}

static jgm_status storedj_apply(jgm_context* ctx, jgm_gemm_precision prec, int transpose,
                                const float* in, size_t nrhs, float* out) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->storedJ_kind == 0) return fail(JGM_ERR_NOT_READY, "no stored J (jgm_storedj_build)");
  if (prec != JGM_GEMM_FP32 && prec != JGM_GEMM_BF16 && prec != JGM_GEMM_BF16_REFINED)
    return fail(JGM_ERR_INVALID_ARGUMENT, "unknown GEMM precision %d", (int)prec);
  if (ctx->storedJ_fp64)
    return fail(JGM_ERR_UNSUPPORTED,
                "the stored J is FP64: its products are jgm_storedj_apply_f64 / _apply_t_f64");
  if (!in || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null operand");
  if (nrhs == 0) return JGM_OK;
  if (nrhs > 4096) return fail(JGM_ERR_UNSUPPORTED, "at most 4096 right-hand sides");
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::launch_storedj_apply(ctx, prec, transpose, in, nrhs, out);
  if (e == hipErrorNotSupported) return fail(JGM_ERR_UNSUPPORTED, "stored J product not available");
  if (e != hipSuccess) return hip_fail(e, "stored J product");
  return JGM_OK;
}

jgm_status jgm_storedj_apply(jgm_context* ctx, jgm_gemm_precision prec, const float* M,
                             size_t nrhs, float* Y) {
  return storedj_apply(ctx, prec, 0, M, nrhs, Y);
}

jgm_status jgm_storedj_apply_t(jgm_context* ctx, jgm_gemm_precision prec, const float* R,
                               size_t nrhs, float* G) {
  return storedj_apply(ctx, prec, 1, R, nrhs, G);
}

static jgm_status storedj_apply_f64(jgm_context* ctx, int transpose, const double* in,
                                    size_t nrhs, double* out) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->storedJ_kind == 0)
    return fail(JGM_ERR_NOT_READY, "no stored J (jgm_storedj_build_fp64)");
  if (!ctx->storedJ_fp64)
    return fail(JGM_ERR_UNSUPPORTED, "the stored J is FP32: build it with jgm_storedj_build_fp64");
  if (!in || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null operand");
  if (nrhs == 0) return JGM_OK;
  if (nrhs > 4096) return fail(JGM_ERR_UNSUPPORTED, "at most 4096 right-hand sides");
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::launch_storedj_apply_f64(ctx, transpose, in, nrhs, out);
  if (e != hipSuccess) return hip_fail(e, "FP64 stored J product");
  return JGM_OK;
}

jgm_status jgm_storedj_apply_f64(jgm_context* ctx, const double* M, size_t nrhs, double* Y) {
  return storedj_apply_f64(ctx, 0, M, nrhs, Y);
}

jgm_status jgm_storedj_apply_t_f64(jgm_context* ctx, const double* R, size_t nrhs, double* G) {
  return storedj_apply_f64(ctx, 1, R, nrhs, G);
}

static jgm_status storedj_ready(jgm_context* ctx, jgm_gemm_precision prec) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (ctx->storedJ_kind == 0) return fail(JGM_ERR_NOT_READY, "no stored J (jgm_storedj_build)");
  if (prec != JGM_GEMM_FP32 && prec != JGM_GEMM_BF16 && prec != JGM_GEMM_BF16_REFINED &&
      prec != JGM_GEMM_FP64)
    return fail(JGM_ERR_INVALID_ARGUMENT, "unknown GEMM precision %d", (int)prec);
  if ((prec == JGM_GEMM_FP64) != ctx->storedJ_fp64)
    return fail(JGM_ERR_UNSUPPORTED,
                ctx->storedJ_fp64 ? "the stored J is FP64: use JGM_GEMM_FP64"
                                  : "JGM_GEMM_FP64 needs the FP64 J of jgm_storedj_build_fp64");
  return check_ready(ctx, ctx->storedJ_kind, JGM_FP64);
}

jgm_status jgm_storedj_forward(jgm_context* ctx, jgm_gemm_precision prec, const double* model,
                               double* out) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (!model || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null model or output pointer");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e = jgm::launch_storedj_forward(ctx, prec, model, out, nullptr);
  if (e != hipSuccess) return hip_fail(e, "stored-J forward");
  return JGM_OK;
}

jgm_status jgm_storedj_gradient(jgm_context* ctx, jgm_gemm_precision prec, const double* model,
                                const double* misfit, double* grad) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (!misfit || !grad) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  if (ctx->transform != JGM_TF_NONE && !model)
    return fail(JGM_ERR_INVALID_ARGUMENT, "a transform needs the model for its Jacobian");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  const int kind = ctx->storedJ_kind;
  const size_t S = ctx->S;
  const size_t kraw = jgm::raw_per_site(kind);
  hipError_t e;
  const double* rm = misfit;
  if (ctx->transform != JGM_TF_NONE) {
    // raw data at the model through the stored J (CachedLQDerivative,
    // FullSensitivityGravMagCalculator.h:170-205)
    if ((e = ctx->raw.ensure(sizeof(double) * S * kraw)) != hipSuccess)
      return hip_fail(e, "stored-J gradient");
    if ((e = ctx->out_stage.ensure(sizeof(double) * S)) != hipSuccess)
      return hip_fail(e, "stored-J gradient");
    if ((e = ctx->misfit_stage.ensure(sizeof(double) * S * kraw)) != hipSuccess)
      return hip_fail(e, "stored-J gradient");
    e = jgm::launch_storedj_forward(ctx, prec, model, ctx->out_stage.as<double>(),
                                    ctx->raw.as<double>());
    if (e != hipSuccess) return hip_fail(e, "stored-J gradient (raw forward)");
    e = jgm::launch_misfit_to_raw(ctx, kind, misfit, ctx->raw.as<const double>(),
                                  ctx->misfit_stage.as<double>());
    if (e != hipSuccess) return hip_fail(e, "stored-J gradient (transform)");
    rm = ctx->misfit_stage.as<const double>();
  }
  e = jgm::launch_storedj_gradient(ctx, prec, rm, grad);
  if (e != hipSuccess) return hip_fail(e, "stored-J gradient");
  return JGM_OK;
}

jgm_status jgm_storedj_forward_batch(jgm_context* ctx, jgm_gemm_precision prec, size_t nmodels,
                                     const double* models, size_t model_ld, double* out,
                                     size_t out_ld) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!models || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null models or output pointer");
  const size_t np = jgm::nparams(ctx, ctx->storedJ_kind),
               nout = ctx->S * data_per_site(ctx, ctx->storedJ_kind);
  if (nmodels > 1 && (model_ld < np || out_ld < nout))
    return fail(JGM_ERR_INVALID_ARGUMENT, "leading dimensions too small");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e = jgm::launch_storedj_forward_batch(ctx, prec, nmodels, models, model_ld, out,
                                                   out_ld, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "stored-J forward batch");
  return JGM_OK;
}

jgm_status jgm_storedj_gradient_batch(jgm_context* ctx, jgm_gemm_precision prec, size_t nmodels,
                                      const double* models, size_t model_ld,
                                      const double* misfits, size_t misfit_ld, double* grads,
                                      size_t grad_ld) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!misfits || !grads) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  if (ctx->transform != JGM_TF_NONE && !models)
    return fail(JGM_ERR_INVALID_ARGUMENT, "a transform needs the models for its Jacobian");
  const int kind = ctx->storedJ_kind;
  const size_t S = ctx->S, kraw = jgm::raw_per_site(kind);
  const size_t np = jgm::nparams(ctx, kind), nmis = S * data_per_site(ctx, kind);
  if (nmodels > 1 && (misfit_ld < nmis || grad_ld < np ||
                      (ctx->transform != JGM_TF_NONE && model_ld < np)))
    return fail(JGM_ERR_INVALID_ARGUMENT, "leading dimensions too small");
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  if ((st = ensure_background(ctx)) != JGM_OK) return st;
  hipError_t e;
  const double* rm = misfits;
  size_t rm_ld = misfit_ld;
  if (ctx->transform != JGM_TF_NONE) {
    const size_t nraw = S * kraw;
    if ((e = ctx->raw.ensure(sizeof(double) * nmodels * nraw)) != hipSuccess ||
        (e = ctx->out_stage.ensure(sizeof(double) * nmodels * S)) != hipSuccess ||
        (e = ctx->misfit_stage.ensure(sizeof(double) * nmodels * nraw)) != hipSuccess)
      return hip_fail(e, "stored-J gradient batch");
    e = jgm::launch_storedj_forward_batch(ctx, prec, nmodels, models, model_ld,
                                          ctx->out_stage.as<double>(), S, ctx->raw.as<double>(),
                                          nraw);
    if (e != hipSuccess) return hip_fail(e, "stored-J gradient batch (raw forward)");
    for (size_t b = 0; b < nmodels && e == hipSuccess; ++b)
      e = jgm::launch_misfit_to_raw(ctx, kind, misfits + b * misfit_ld,
                                    ctx->raw.as<const double>() + b * nraw,
                                    ctx->misfit_stage.as<double>() + b * nraw);
    if (e != hipSuccess) return hip_fail(e, "stored-J gradient batch (transform)");
    rm = ctx->misfit_stage.as<const double>();
    rm_ld = nraw;
  }
  e = jgm::launch_storedj_gradient_batch(ctx, prec, nmodels, rm, rm_ld, grads, grad_ld);
  if (e != hipSuccess) return hip_fail(e, "stored-J gradient batch");
  return JGM_OK;
}

jgm_status jgm_storedj_forward_batch_host(jgm_context* ctx, jgm_gemm_precision prec,
                                          size_t nmodels, const double* models, double* out) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!models || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null models or output pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, ctx->storedJ_kind),
               nout = ctx->S * data_per_site(ctx, ctx->storedJ_kind);
  const HostArg args[2] = {{models, nullptr, nmodels * np}, {nullptr, out, nmodels * nout}};
  return run_staged(ctx, "storedj_forward_batch_host", args,
                    [&](const double* const* in, double* const* o) {
                      return jgm_storedj_forward_batch(ctx, prec, nmodels, in[0], np, o[1], nout);
                    });
}

jgm_status jgm_storedj_gradient_batch_host(jgm_context* ctx, jgm_gemm_precision prec,
                                           size_t nmodels, const double* models,
                                           const double* misfits, double* grads) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (nmodels == 0) return JGM_OK;
  if (!misfits || !grads) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, ctx->storedJ_kind),
               nmis = ctx->S * data_per_site(ctx, ctx->storedJ_kind);
  const HostArg args[3] = {{models, nullptr, models ? nmodels * np : 0},
                           {misfits, nullptr, nmodels * nmis},
                           {nullptr, grads, nmodels * np}};
  return run_staged(ctx, "storedj_gradient_batch_host", args,
                    [&](const double* const* in, double* const* o) {
                      return jgm_storedj_gradient_batch(ctx, prec, nmodels,
                                                        models ? in[0] : nullptr, np, in[1], nmis,
                                                        o[2], np);
                    });
}

jgm_status jgm_storedj_forward_host(jgm_context* ctx, jgm_gemm_precision prec,
                                    const double* model, double* out) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (!model || !out) return fail(JGM_ERR_INVALID_ARGUMENT, "null model or output pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, ctx->storedJ_kind);
  const size_t nout = ctx->S * data_per_site(ctx, ctx->storedJ_kind);
  hipError_t e;
  if ((e = ctx->model_stage.ensure(sizeof(double) * (np ? np : 1))) != hipSuccess)
    return hip_fail(e, "storedj_forward_host");
  if ((e = ctx->grad_stage.ensure(sizeof(double) * (nout ? nout : 1))) != hipSuccess)
    return hip_fail(e, "storedj_forward_host");
  if ((e = hipMemcpyAsync(ctx->model_stage.ptr, model, sizeof(double) * np,
                          hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
    return hip_fail(e, "storedj_forward_host");
  st = jgm_storedj_forward(ctx, prec, ctx->model_stage.as<double>(), ctx->grad_stage.as<double>());
  if (st != JGM_OK) return st;
  if (nout && (e = hipMemcpyAsync(out, ctx->grad_stage.ptr, sizeof(double) * nout,
                                  hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
    return hip_fail(e, "storedj_forward_host");
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return hip_fail(e, "storedj_forward_host");
  return JGM_OK;
}

jgm_status jgm_storedj_gradient_host(jgm_context* ctx, jgm_gemm_precision prec,
                                     const double* model, const double* misfit, double* grad) {
  jgm_status st = storedj_ready(ctx, prec);
  if (st != JGM_OK) return st;
  if (!misfit || !grad) return fail(JGM_ERR_INVALID_ARGUMENT, "null misfit or gradient pointer");
  DeviceGuard g(ctx->device);
  const size_t np = jgm::nparams(ctx, ctx->storedJ_kind);
  const size_t nmis = ctx->S * data_per_site(ctx, ctx->storedJ_kind);
  jgm::DeviceBuffer dmis;  // separate from the internal misfit_stage used by transforms
  hipError_t e = hipSuccess;
  if ((e = ctx->model_stage.ensure(sizeof(double) * (np ? np : 1))) != hipSuccess)
    return hip_fail(e, "storedj_gradient_host");
  if ((e = ctx->grad_stage.ensure(sizeof(double) * (np ? np : 1))) != hipSuccess)
    return hip_fail(e, "storedj_gradient_host");
  if ((e = dmis.ensure(sizeof(double) * (nmis ? nmis : 1))) != hipSuccess)
    return hip_fail(e, "storedj_gradient_host");
  if (model && (e = hipMemcpyAsync(ctx->model_stage.ptr, model, sizeof(double) * np,
                                   hipMemcpyHostToDevice, ctx->stream)) != hipSuccess) {
    dmis.release();
    return hip_fail(e, "storedj_gradient_host");
  }
  if (nmis && (e = hipMemcpyAsync(dmis.ptr, misfit, sizeof(double) * nmis, hipMemcpyHostToDevice,
                                  ctx->stream)) != hipSuccess) {
    dmis.release();
    return hip_fail(e, "storedj_gradient_host");
  }
  st = jgm_storedj_gradient(ctx, prec, model ? ctx->model_stage.as<double>() : nullptr,
                            dmis.as<double>(), ctx->grad_stage.as<double>());
  if (st == JGM_OK)
    e = hipMemcpyAsync(grad, ctx->grad_stage.ptr, sizeof(double) * np, hipMemcpyDeviceToHost,
                       ctx->stream);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  dmis.release();
  if (st != JGM_OK) return st;
  if (e != hipSuccess) return hip_fail(e, "storedj_gradient_host");
  if (e2 != hipSuccess) return hip_fail(e2, "storedj_gradient_host");
  return JGM_OK;
}

jgm_status jgm_storedj_release(jgm_context* ctx) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->storedJ.release();
  ctx->storedJ_bf16.release();
  ctx->storedJ_bf16c.release();
  ++ctx->storedJ_gen;
  jgm::release_host_storedj(ctx);
  ctx->acc64.release();
  ctx->storedJ_kind = 0;
  ctx->storedJ_fp64 = false;
  ctx->storedJ_rows = ctx->storedJ_cols = 0;
  return JGM_OK;
}

jgm_status jgm_device_alloc(jgm_context* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr) return fail(JGM_ERR_INVALID_ARGUMENT, "null context or output pointer");
  *ptr = nullptr;
  DeviceGuard g(ctx->device);
  if (g.err != hipSuccess) return hip_fail(g.err, "hipSetDevice");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return hip_fail(e, "jgm_device_alloc");
  return JGM_OK;
}

jgm_status jgm_device_free(jgm_context* ctx, void* ptr) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (!ptr) return JGM_OK;
  DeviceGuard g(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  hipError_t e = hipFree(ptr);
  if (e != hipSuccess) return hip_fail(e, "jgm_device_free");
  return JGM_OK;
}

static jgm_status copy(jgm_context* ctx, void* dst, const void* src, size_t bytes,
                       hipMemcpyKind kind, const char* what) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  if (bytes == 0) return JGM_OK;
  if (!dst || !src) return fail(JGM_ERR_INVALID_ARGUMENT, "%s: null pointer", what);
  DeviceGuard g(ctx->device);
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(e, what);
  return JGM_OK;
}

jgm_status jgm_copy_to_device(jgm_context* ctx, void* dst, const void* src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyHostToDevice, "jgm_copy_to_device");
}

jgm_status jgm_copy_to_host(jgm_context* ctx, void* dst, const void* src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost, "jgm_copy_to_host");
}

jgm_status jgm_copy_device(jgm_context* ctx, void* dst, const void* src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyDeviceToDevice, "jgm_copy_device");
}

jgm_status jgm_profile_enable(jgm_context* ctx, int enable) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  ctx->prof.enabled = enable != 0;
  return JGM_OK;
}

jgm_status jgm_profile_reset(jgm_context* ctx) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  DeviceGuard g(ctx->device);
  jgm::prof_clear(ctx);
  return JGM_OK;
}

size_t jgm_profile_count(jgm_context* ctx) {
  if (!ctx) return 0;
  DeviceGuard g(ctx->device);
  (void)jgm::prof_harvest(ctx);
  return ctx->prof.entries.size();
}

jgm_status jgm_profile_get(jgm_context* ctx, size_t index, const char** name, double* total_ms,
                           size_t* launches) {
  if (!ctx) return fail(JGM_ERR_INVALID_ARGUMENT, "null context");
  DeviceGuard g(ctx->device);
  hipError_t e = jgm::prof_harvest(ctx);
  if (e != hipSuccess) return hip_fail(e, "profile");
  if (index >= ctx->prof.entries.size())
    return fail(JGM_ERR_INVALID_ARGUMENT, "profile index %zu out of range", index);
  const auto& en = ctx->prof.entries[index];
  if (name) *name = en.name.c_str();
  if (total_ms) *total_ms = en.total_ms;
  if (launches) *launches = en.launches;
  return JGM_OK;
}

}  // extern "C"
