// A-posteriori check of the mixed-precision path.  The mixed path states its error
// relative to each component's largest magnitude over the survey (forward) or to the
// largest gradient entry (adjoint).  That bound holds when the cells' contributions do not
// cancel much; on models whose field is a small remainder of large cell fields (a nearly
// uniform magnetised slab) the far field's FP32 error per term is amplified.  After a
// mixed call the FP64 closed form is evaluated at a few sample sites (forward: every
// corner of the mesh for each sample site, corner-parallel) or sample cells (adjoint:
// every site for each sample cell) and compared with the mixed result; when the sampled
// error exceeds the tolerance the whole call is recomputed in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "epilogue.hpp"
#include "prism_device.hpp"

namespace jgm {

namespace {

constexpr int kBlock = 256;
inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Sum over the block (256 threads = 4 waves): within each wave by lane shuffles, then
// the four wave sums in a fixed order.  Every thread gets the result.
__device__ inline double block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const double r = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return r;
}

// w(I,J,K) = sum sigma rho over the eight cells sharing the corner (rho zero outside)
__global__ __launch_bounds__(kBlock) void k_check_weights(const double* __restrict__ rho,
                                                          int nx, int ny, int nz,
                                                          double* __restrict__ w) {
  const long ncx = nx + 1, ncy = ny + 1, ncz = nz + 1;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncx * ncy * ncz) return;
  const int K = (int)(q % ncz);
  const long t = q / ncz;
  const int J = (int)(t % ncy);
  const int I = (int)(t / ncy);
  double acc = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int i = I - a, j = J - b, k = K - c;
        if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) continue;
        const double v = rho[((long)i * ny + j) * nz + k];
        acc += ((a + b + c) & 1) ? v : -v;
      }
  w[q] = acc;
}

// Forward at F sample sites: block (x, f) sums w(q) F_m(s_f, q) over a strided share of
// the corners; partial[(x F + f) NC + m].
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_check_forward(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const int* __restrict__ sidx, int F, const double* __restrict__ Xc,
    const double* __restrict__ Yc, const double* __restrict__ Zc, int ncy, int ncz, long nq,
    const double* __restrict__ w, double* __restrict__ partial) {
  constexpr int NC = Ncomp<MASK>::value;
  __shared__ double red[kBlock / 64];
  const long stride = (long)gridDim.x * blockDim.x;
  const int f = blockIdx.y;
  const int s = sidx[f];
  const double px = sx[s], py = sy[s], pz = sz[s];
  double acc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = 0.0;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += stride) {
    const int K = (int)(q % ncz);
    const long t = q / ncz;
    const int J = (int)(t % ncy);
    const int I = (int)(t / ncy);
    double tv[NC];
    corner_terms_f64<MASK>(px - Xc[I], py - Yc[J], pz - Zc[K], tv);
    const double wq = w[q];
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[m] = fma(wq, tv[m], acc[m]);
  }
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    const double v = block_sum(acc[m], red);
    if (threadIdx.x == 0) partial[((long)blockIdx.x * F + f) * NC + m] = v;
  }
}

// Survey maximum of |raw| per raw component: block m writes smax[m]
__global__ __launch_bounds__(kBlock) void k_check_survey_max(const double* __restrict__ raw,
                                                             int S, int kraw,
                                                             double* __restrict__ smax) {
  __shared__ double red[kBlock];
  const int m = blockIdx.x;
  double v = 0.0;
  for (int s = threadIdx.x; s < S; s += blockDim.x) v = fmax(v, fabs(raw[(long)s * kraw + m]));
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) smax[m] = red[0];
}

// FP64 raw data of sample site f (block f: the block partials summed as a tree, then
// background and assembly as the forward) against the mixed raw data;
// errs[f] = largest |difference| / survey maximum of its component.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_check_forward_sites(
    const double* __restrict__ partial, int nblocks, const int* __restrict__ sidx, int F,
    const double* __restrict__ bgT, const double* __restrict__ bgvals, int L, int kind,
    FieldDir fd, const double* __restrict__ mixed_raw, const double* __restrict__ smax,
    double* __restrict__ errs) {
  constexpr int NC = Ncomp<MASK>::value;
  __shared__ double red[kBlock / 64];
  const int f = blockIdx.x;
  double acc[NC], sc[NC];
  comp_scales<MASK>(sc);
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    double a = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) a += partial[((long)b * F + f) * NC + m];
    acc[m] = block_sum(a, red) * sc[m];
  }
  if (threadIdx.x != 0) return;
  double u[7], raw[10];
  spread_components<MASK>(acc, u);
  const int s = sidx[f];
  const int kraw = assemble_raw(u, s, bgT, bgvals, L, kind, fd, raw);
  double e = 0.0;
  for (int m = 0; m < kraw; ++m) {
    const double d = fabs(raw[m] - mixed_raw[(long)s * kraw + m]);
    e = fmax(e, d / fmax(smax[m], 1e-300));
  }
  errs[f] = e;
}

// out[0] = max of v[0..n).  One block.
__global__ __launch_bounds__(kBlock) void k_check_max(const double* __restrict__ v, int n,
                                                      double* __restrict__ out) {
  __shared__ double red[kBlock];
  double e = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) e = fmax(e, v[i]);
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// pm[block] = max of |g| over a grid-strided share of g[0..n)
__global__ __launch_bounds__(kBlock) void k_check_absmax(const double* __restrict__ g, long n,
                                                         double* __restrict__ pm) {
  __shared__ double red[kBlock];
  double e = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    e = fmax(e, fabs(g[i]));
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) pm[blockIdx.x] = red[0];
}

// Adjoint at FC sample cells: blocks (cell, split), threads over the split's sites,
// g_c = 2 sum_s sum_m c_m(s) J_m(s, c) with the FP64 box term of the cell (the gridded
// part; the background gradient is exact in both paths); gs[cell gridDim.y + split].
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_check_adjoint(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const double* __restrict__ rm, int kind, FieldDir fd, const int* __restrict__ cidx,
    const double* __restrict__ xlo, const double* __restrict__ ylo,
    const double* __restrict__ zlo, const double* __restrict__ xsz,
    const double* __restrict__ ysz, const double* __restrict__ zsz, int ny, int nz,
    double* __restrict__ gs) {
  constexpr int NC = Ncomp<MASK>::value;
  __shared__ double red[kBlock];
  const long c = cidx[blockIdx.x];
  const int k = (int)(c % nz);
  const long t = c / nz;
  const int j = (int)(t % ny);
  const int i = (int)(t / ny);
  double acc = 0.0;
  for (int s = blockIdx.y * blockDim.x + threadIdx.x; s < S; s += gridDim.y * blockDim.x) {
    double v[NC], u[7], c7[7];
    box_terms_f64<MASK>(sx[s], sy[s], sz[s], xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
    spread_components<MASK>(v, u);
    site_coefficients(rm, s, kind, fd, c7);
    for (int m = 0; m < 7; ++m) acc = fma(c7[m], u[m], acc);
  }
  const double g = block_sum(acc, red);
  if (threadIdx.x == 0) gs[(long)blockIdx.x * gridDim.y + blockIdx.y] = 2.0 * g;
}

// err[0] = max over sample cells |g_mixed - g_fp64| / max over cells |g_mixed| (the latter
// from the nb block maxima pm).  One block.
__global__ __launch_bounds__(kBlock) void k_check_adjoint_finish(const double* __restrict__ grad,
                                                                 const double* __restrict__ pm,
                                                                 int nb,
                                                                 const int* __restrict__ cidx,
                                                                 const double* __restrict__ gs,
                                                                 int FC, int nsplit,
                                                                 double* __restrict__ err) {
  __shared__ double red[kBlock];
  double gmax = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) gmax = fmax(gmax, pm[b]);
  red[threadIdx.x] = gmax;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  gmax = red[0];
  __syncthreads();
  double e = 0.0;
  for (int f = threadIdx.x; f < FC; f += blockDim.x) {
    double g = 0.0;
    for (int q = 0; q < nsplit; ++q) g += gs[(long)f * nsplit + q];
    e = fmax(e, fabs(grad[cidx[f]] - g) / fmax(gmax, 1e-300));
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) err[0] = red[0];
}

int check_mask(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return kMaskScalar;
    case JGM_KIND_SCALAR_TENSOR: return kMaskAll;
    default: return kMaskTensor;
  }
}

// Sample sites: evenly spaced in survey order; sample cells: evenly spaced in cell order,
// offset by half a stride.  Uploaded once per geometry and sample size.
hipError_t ensure_samples(jgm_context* ctx) {
  const size_t S = ctx->S, N = ncells(ctx);
  const size_t F = std::min(ctx->check_sites, S), FC = std::min(ctx->check_cells, N);
  if (ctx->check_gen == ctx->config_gen && ctx->check_F == F && ctx->check_FC == FC)
    return hipSuccess;
  std::vector<int> si(std::max<size_t>(F, 1)), ci(std::max<size_t>(FC, 1));
  for (size_t f = 0; f < F; ++f) si[f] = (int)((2 * f + 1) * S / (2 * F));
  for (size_t f = 0; f < FC; ++f) ci[f] = (int)((2 * f + 1) * N / (2 * FC));
  hipError_t e;
  if ((e = ctx->check_sidx.ensure(sizeof(int) * si.size())) != hipSuccess ||
      (e = ctx->check_cidx.ensure(sizeof(int) * ci.size())) != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(ctx->check_sidx.ptr, si.data(), sizeof(int) * si.size(),
                          hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(ctx->check_cidx.ptr, ci.data(), sizeof(int) * ci.size(),
                          hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
    return e;
  ctx->check_F = F;
  ctx->check_FC = FC;
  ctx->check_gen = ctx->config_gen;
  return hipSuccess;
}

template <int MASK>
hipError_t check_forward_impl(jgm_context* ctx, int kind, const double* model,
                              const double* mixed_raw, double* err_host) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S, F = (int)ctx->check_F;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const int kraw = (int)raw_per_site(kind);
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long nq = (long)ncorners(ctx);
  // blocks per sample site: about 2 CUs' worth over all sites, each thread taking at least
  // 8 corners so that the block sums stay a small part of the work (on small meshes a floor
  // of 32 left most CUs idle: config 1's 139K corners gave 17 blocks per site)
  const long want = std::max<long>(1, 2L * ctx->cu_count / std::max(1, F));
  const int nblocks = (int)std::max<long>(1, std::min<long>(want, (nq + 8L * kBlock - 1) / (8L * kBlock)));
  hipError_t e;
  // partials (nblocks F NC), survey maxima (kraw), per-site errors (F), the error (1)
  if ((e = ctx->check_w.ensure(sizeof(double) * (size_t)nq)) != hipSuccess ||
      (e = ctx->check_part.ensure(sizeof(double) * ((size_t)nblocks * F * NC + kraw + F + 1))) !=
          hipSuccess)
    return e;
  double* part = ctx->check_part.as<double>();
  double* smax = part + (size_t)nblocks * F * NC;
  double* errs = smax + kraw;
  double* err = errs + F;
  {
    ProfScope ps(ctx, "mixed_check_forward_fp64");
    hipLaunchKernelGGL(k_check_weights, dim3(blocks_for(nq)), dim3(kBlock), 0, st, model,
                       (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, ctx->check_w.as<double>());
    hipLaunchKernelGGL((k_check_forward<MASK>), dim3(nblocks, F), dim3(kBlock), 0, st,
                       ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                       ctx->d_sz.as<const double>(), ctx->check_sidx.as<const int>(), F,
                       ctx->d_Xc.as<const double>(), ctx->d_Yc.as<const double>(),
                       ctx->d_Zc.as<const double>(), ncy, ncz, nq,
                       ctx->check_w.as<const double>(), part);
    hipLaunchKernelGGL(k_check_survey_max, dim3(kraw), dim3(kBlock), 0, st, mixed_raw, S, kraw,
                       smax);
    hipLaunchKernelGGL((k_check_forward_sites<MASK>), dim3(F), dim3(kBlock), 0, st, part,
                       nblocks, ctx->check_sidx.as<const int>(), F, ctx->bgT.as<const double>(),
                       model + N, L, kind, field_dir(ctx), mixed_raw, smax, errs);
    hipLaunchKernelGGL(k_check_max, dim3(1), dim3(kBlock), 0, st, errs, F, err);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(err_host, err, sizeof(double), hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return e;
  return hipStreamSynchronize(st);
}

template <int MASK>
hipError_t check_adjoint_impl(jgm_context* ctx, int kind, const double* raw_misfit,
                              const double* grad, double* err_host) {
  hipStream_t st = ctx->stream;
  const int FC = (int)ctx->check_FC;
  const long N = (long)ncells(ctx);
  const int nsplit = (int)std::max<size_t>(1, std::min<size_t>(32, (ctx->S + kBlock - 1) / kBlock));
  const int nb = (int)std::max<long>(1, std::min<long>(2L * ctx->cu_count, (N + kBlock - 1) / kBlock));
  hipError_t e;
  // per-split sums (FC nsplit), block maxima of |grad| (nb), the error (1)
  if ((e = ctx->check_part.ensure(sizeof(double) * ((size_t)FC * nsplit + nb + 1))) != hipSuccess)
    return e;
  double* gs = ctx->check_part.as<double>();
  double* pm = gs + (size_t)FC * nsplit;
  double* err = pm + nb;
  {
    ProfScope ps(ctx, "mixed_check_adjoint_fp64");
    hipLaunchKernelGGL((k_check_adjoint<MASK>), dim3(FC, nsplit), dim3(kBlock), 0, st,
                       ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                       ctx->d_sz.as<const double>(), (int)ctx->S, raw_misfit, kind, field_dir(ctx),
                       ctx->check_cidx.as<const int>(), ctx->d_xlo.as<const double>(),
                       ctx->d_ylo.as<const double>(), ctx->d_zlo.as<const double>(),
                       ctx->d_xsz.as<const double>(), ctx->d_ysz.as<const double>(),
                       ctx->d_zsz.as<const double>(), (int)ctx->ny, (int)ctx->nz, gs);
    hipLaunchKernelGGL(k_check_absmax, dim3(nb), dim3(kBlock), 0, st, grad, N, pm);
    hipLaunchKernelGGL(k_check_adjoint_finish, dim3(1), dim3(kBlock), 0, st, grad, pm, nb,
                       ctx->check_cidx.as<const int>(), gs, FC, nsplit, err);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(err_host, err, sizeof(double), hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return e;
  return hipStreamSynchronize(st);
}

}  // namespace

bool mixed_check_applies(const jgm_context* ctx, int kind) {
  return ctx->check_sites > 0 && ncells(ctx) > 0 && ctx->S > 0 &&
         (kind == JGM_KIND_SCALAR || kind == JGM_KIND_TENSOR || kind == JGM_KIND_MAGNETIC ||
          kind == JGM_KIND_SCALAR_TENSOR);
}

hipError_t launch_mixed_check_forward(jgm_context* ctx, int kind, const double* model,
                                      const double* mixed_raw, double* err_host) {
  hipError_t e = ensure_samples(ctx);
  if (e != hipSuccess) return e;
  switch (check_mask(kind)) {
    case kMaskScalar: return check_forward_impl<kMaskScalar>(ctx, kind, model, mixed_raw, err_host);
    case kMaskAll: return check_forward_impl<kMaskAll>(ctx, kind, model, mixed_raw, err_host);
    default: return check_forward_impl<kMaskTensor>(ctx, kind, model, mixed_raw, err_host);
  }
}

hipError_t launch_mixed_check_adjoint(jgm_context* ctx, int kind, const double* raw_misfit,
                                      const double* grad, double* err_host) {
  hipError_t e = ensure_samples(ctx);
  if (e != hipSuccess) return e;
  if (ctx->check_FC == 0) {
    *err_host = 0.0;
    return hipSuccess;
  }
  switch (check_mask(kind)) {
    case kMaskScalar:
      return check_adjoint_impl<kMaskScalar>(ctx, kind, raw_misfit, grad, err_host);
    case kMaskAll: return check_adjoint_impl<kMaskAll>(ctx, kind, raw_misfit, grad, err_host);
    default: return check_adjoint_impl<kMaskTensor>(ctx, kind, raw_misfit, grad, err_host);
  }
}

}  // namespace jgm
