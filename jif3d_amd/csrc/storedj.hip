// Stored-sensitivity path: J built once in HBM (FP32), then applied to many
// right-hand sides at once on the matrix cores.
//
// J layout: column-major, J[c * R + row], rows = site * kJ + component fastest, with
// the compact components per site: gz (scalar); Uxx, Uxy, Uxz, Uyy, Uyz, Uzz
// (tensor); Bx, By, Bz (magnetic); gz then the six tensor values (scalar+tensor).
// Columns are the mesh cells (background layers are not stored).
//
// Products (nrhs right-hand sides, column-major operands):
//   Y = J M    (R x nrhs):  split over cells; A fragments of J are read straight from
//                          HBM (32 consecutive rows per cell = one 128-byte line).
//   G = J^T Rr (N x nrhs):  one 128-cell tile per workgroup over all rows; J tiles are
//                          staged through LDS so global reads stay coalesced.
// M and Rr are first transposed to [k][nrhs_padded] (FP32, nrhs padded to 64).
// Precisions: FP32 operands on v_mfma_f32_32x32x2_f32; BF16 operands on
// v_mfma_f32_32x32x16_bf16 (FP32 accumulate), with the "refined" variant splitting
// both operands into hi + lo BF16 parts in registers and summing hi*hi + hi*lo +
// lo*hi (per-product error <= 3 * 2^-16 relative instead of 2^-7).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "context.hpp"
#include "epilogue.hpp"
#include "prism_device.hpp"

namespace jgm {

namespace {

constexpr int kBlock = 256;
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__host__ __device__ inline int compact_rows(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return 1;
    case JGM_KIND_TENSOR: return 6;
    case JGM_KIND_MAGNETIC: return 3;
    default: return 7;
  }
}

int mask_of(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return kMaskScalar;
    case JGM_KIND_SCALAR_TENSOR: return kMaskAll;
    default: return kMaskTensor;
  }
}

// Build: one lane per site, the wave walks a chunk of cells (wave-uniform); FP64
// closed form per cell (8 corners), stored as FP32 in the compact layout.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_storedj_build(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const double* __restrict__ xlo, const double* __restrict__ ylo,
    const double* __restrict__ zlo, const double* __restrict__ xsz,
    const double* __restrict__ ysz, const double* __restrict__ zsz, int ny, int nz, long ncell,
    long cells_per_chunk, int kind, FieldDir fd, float* __restrict__ J) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = s < S ? s : S - 1;
  const double px = sx[sl], py = sy[sl], pz = sz[sl];
  const int kJ = compact_rows(kind);
  const long R = (long)S * kJ;
  const long c0 = (long)blockIdx.y * cells_per_chunk;
  const long c1 = c0 + cells_per_chunk < ncell ? c0 + cells_per_chunk : ncell;
  for (long c = c0; c < c1; ++c) {
    const int k = (int)(c % nz);
    const long t = c / nz;
    const int j = (int)(t % ny);
    const int i = (int)(t / ny);
    double v[NC], u[7];
    box_terms_f64<MASK>(px, py, pz, xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
    spread_components<MASK>(v, u);  // gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz
    if (s >= S) continue;
    float* o = J + c * R + (long)s * kJ;
    if (kind == JGM_KIND_SCALAR) {
      o[0] = (float)u[0];
    } else if (kind == JGM_KIND_MAGNETIC) {
      o[0] = (float)((u[1] * fd.bx + u[4] * fd.by + u[5] * fd.bz) * fd.scale);
      o[1] = (float)((u[4] * fd.bx + u[2] * fd.by + u[6] * fd.bz) * fd.scale);
      o[2] = (float)((u[5] * fd.bx + u[6] * fd.by + u[3] * fd.bz) * fd.scale);
    } else {
      int q = 0;
      if (kind == JGM_KIND_SCALAR_TENSOR) o[q++] = (float)u[0];
      o[q + 0] = (float)u[1];
      o[q + 1] = (float)u[4];
      o[q + 2] = (float)u[5];
      o[q + 3] = (float)u[2];
      o[q + 4] = (float)u[6];
      o[q + 5] = (float)u[3];
    }
  }
}

// Column-major X (n x nrhs) -> Xt[k][P] (P = nrhs padded to a multiple of 64, zeros).
__global__ __launch_bounds__(kBlock) void k_transpose_pad(const float* __restrict__ X, long n,
                                                          int nrhs, int P,
                                                          float* __restrict__ Xt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * P) return;
  const long k = idx / P;
  const int b = (int)(idx % P);
  Xt[idx] = b < nrhs ? X[k + (long)b * n] : 0.0f;
}

// x = hi + lo + e, |e| <= 2^-16 |x|: hi is x rounded to BF16, lo the rounded remainder
// (x - hi is exact in FP32)
__device__ __forceinline__ void split_bf16(float x, bf16x8& hi, bf16x8& lo, int e) {
  const __bf16 h = (__bf16)x;
  hi[e] = h;
  lo[e] = (__bf16)(x - (float)h);
}

// Y partials: workgroup = 4 waves x 32 rows, 64 right-hand sides (two 32x32 tiles),
// cells [c0, c1) of one split.  Yp[split][row][P].
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_gemm_jm(const float* __restrict__ J, long R, long N,
                                                    const float* __restrict__ Mt, int P,
                                                    long cells_per_split,
                                                    float* __restrict__ Yp) {
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const long row0 = ((long)blockIdx.x * 4 + wave) * 32;
  const int b0 = blockIdx.z * 64;
  const long c0 = (long)blockIdx.y * cells_per_split;
  const long c1 = c0 + cells_per_split < N ? c0 + cells_per_split : N;
  const int li = lane & 31, lk = lane >> 5;
  const long row = row0 + li;
  const bool rv = row < R;
  f32x16 acc0 = {0}, acc1 = {0};
  if constexpr (PREC == JGM_GEMM_FP32) {
    // every lane joins every MFMA: the k-slot past the end of the range reads zeros
    for (long c = c0; c < c1; c += 2) {
      const long cc = c + lk;
      const bool cv = cc < c1;
      const float a = (rv && cv) ? J[cc * R + row] : 0.0f;
      const float bv0 = cv ? Mt[cc * P + b0 + li] : 0.0f;
      const float bv1 = cv ? Mt[cc * P + b0 + 32 + li] : 0.0f;
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv1, acc1, 0, 0, 0);
    }
  } else {
    for (long c = c0; c < c1; c += 16) {
      bf16x8 ahi, alo, bhi0, blo0, bhi1, blo1;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const long cc = c + 8 * lk + e;
        const bool cv = cc < c1;
        const float a = (rv && cv) ? J[cc * R + row] : 0.0f;
        const float m0 = cv ? Mt[cc * P + b0 + li] : 0.0f;
        const float m1 = cv ? Mt[cc * P + b0 + 32 + li] : 0.0f;
        if constexpr (PREC == JGM_GEMM_BF16) {
          ahi[e] = (__bf16)a;
          bhi0[e] = (__bf16)m0;
          bhi1[e] = (__bf16)m1;
        } else {
          split_bf16(a, ahi, alo, e);
          split_bf16(m0, bhi0, blo0, e);
          split_bf16(m1, bhi1, blo1, e);
        }
      }
      acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi1, acc1, 0, 0, 0);
      if constexpr (PREC == JGM_GEMM_BF16_REFINED) {
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo1, acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi1, acc1, 0, 0, 0);
      }
    }
  }
  // D layout of a 32x32 MFMA: register r of lane l holds (row 8*(r/4) + 4*(l/32) + r%4, col l%32)
  float* out = Yp + (long)blockIdx.y * R * P;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long orow = row0 + 8 * (r >> 2) + 4 * lk + (r & 3);
    if (orow < R) {
      out[orow * P + b0 + li] = acc0[r];
      out[orow * P + b0 + 32 + li] = acc1[r];
    }
  }
}

// Y[row + b*R] = sum over splits of Yp[split][row][b]
__global__ __launch_bounds__(kBlock) void k_gemm_jm_reduce(const float* __restrict__ Yp,
                                                           int nsplit, long R, int P, int nrhs,
                                                           float* __restrict__ Y) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= R * nrhs) return;
  const long row = idx / nrhs;
  const int b = (int)(idx % nrhs);
  double a = 0.0;
  for (int s = 0; s < nsplit; ++s) a += Yp[(long)s * R * P + row * P + b];
  Y[row + (long)b * R] = (float)a;
}

// G = J^T Rt: workgroup = 128 cells (4 waves x 32) x 64 right-hand sides, all rows.
// J tiles [128 cells][32 rows] pass through LDS (coalesced 128-byte reads per cell).
template <int PREC>
__global__ __launch_bounds__(kBlock) void k_gemm_jtr(const float* __restrict__ J, long R, long N,
                                                     const float* __restrict__ Rt, int P,
                                                     float* __restrict__ G, int nrhs) {
  constexpr int TR = 32;        // rows per stage
  constexpr int PITCH = TR + 1;  // LDS row pitch (floats) against bank conflicts
  __shared__ float tile[128 * PITCH];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const long cbase = (long)blockIdx.x * 128;
  const int b0 = blockIdx.y * 64;
  const int li = lane & 31, lk = lane >> 5;
  f32x16 acc0 = {0}, acc1 = {0};
  for (long r0 = 0; r0 < R; r0 += TR) {
    __syncthreads();
    // 32 consecutive lanes read the 32 rows (one 128-byte line) of one cell
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int idx = e * kBlock + threadIdx.x;
      const int lc = idx >> 5, lr = idx & 31;
      const long cell = cbase + lc, rr = r0 + lr;
      tile[lc * PITCH + lr] = (cell < N && rr < R) ? J[cell * R + rr] : 0.0f;
    }
    __syncthreads();
    const int ci = wave * 32 + li;  // this lane's cell row of the A fragment
    if constexpr (PREC == JGM_GEMM_FP32) {
#pragma unroll
      for (int ks = 0; ks < TR; ks += 2) {
        const float a = tile[ci * PITCH + ks + lk];
        const long rr = r0 + ks + lk;
        const float bv0 = rr < R ? Rt[rr * P + b0 + li] : 0.0f;
        const float bv1 = rr < R ? Rt[rr * P + b0 + 32 + li] : 0.0f;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv1, acc1, 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < TR; ks += 16) {
        bf16x8 ahi, alo, bhi0, blo0, bhi1, blo1;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int kk = ks + 8 * lk + e;
          const long rr = r0 + kk;
          const float a = tile[ci * PITCH + kk];
          const float m0 = rr < R ? Rt[rr * P + b0 + li] : 0.0f;
          const float m1 = rr < R ? Rt[rr * P + b0 + 32 + li] : 0.0f;
          if constexpr (PREC == JGM_GEMM_BF16) {
            ahi[e] = (__bf16)a;
            bhi0[e] = (__bf16)m0;
            bhi1[e] = (__bf16)m1;
          } else {
            split_bf16(a, ahi, alo, e);
            split_bf16(m0, bhi0, blo0, e);
            split_bf16(m1, bhi1, blo1, e);
          }
        }
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi1, acc1, 0, 0, 0);
        if constexpr (PREC == JGM_GEMM_BF16_REFINED) {
          acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo1, acc1, 0, 0, 0);
          acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi1, acc1, 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const long c = cbase + wave * 32 + 8 * (r >> 2) + 4 * lk + (r & 3);
    if (c < N) {
      if (b0 + li < nrhs) G[c + (long)(b0 + li) * N] = acc0[r];
      if (b0 + 32 + li < nrhs) G[c + (long)(b0 + 32 + li) * N] = acc1[r];
    }
  }
}

}  // namespace

hipError_t launch_storedj_build(jgm_context* ctx, int kind, int precision) {
  (void)precision;  // the build evaluates the FP64 closed form; J is stored in FP32
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int kJ = compact_rows(kind);
  const long R = (long)S * kJ;
  if (S == 0 || N == 0) return hipErrorInvalidValue;
  hipError_t e = ctx->storedJ.ensure(sizeof(float) * (size_t)R * N);
  if (e != hipSuccess) return e;
  const int site_blocks = (S + kBlock - 1) / kBlock;
  const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
  long want = std::max<long>(1, std::min<long>(N, (16384 + waves_per_chunk - 1) / waves_per_chunk));
  const long per = (N + want - 1) / want;
  const int nchunks = (int)((N + per - 1) / per);
  const int ny = (int)ctx->ny, nz = (int)ctx->nz;
  FieldDir fd = field_dir(ctx);
  {
    ProfScope ps(ctx, "storedj_build_fp64");
#define JGM_BUILD(M)                                                                            \
  hipLaunchKernelGGL((k_storedj_build<M>), dim3(site_blocks, nchunks), dim3(kBlock), 0,        \
                     ctx->stream, ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),               \
                     ctx->d_sz.as<double>(), S, ctx->d_xlo.as<double>(), ctx->d_ylo.as<double>(), \
                     ctx->d_zlo.as<double>(), ctx->d_xsz.as<double>(), ctx->d_ysz.as<double>(), \
                     ctx->d_zsz.as<double>(), ny, nz, N, per, kind, fd, ctx->storedJ.as<float>())
    switch (mask_of(kind)) {
      case kMaskScalar: JGM_BUILD(kMaskScalar); break;
      case kMaskAll: JGM_BUILD(kMaskAll); break;
      default: JGM_BUILD(kMaskTensor); break;
    }
#undef JGM_BUILD
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  ctx->storedJ_kind = kind;
  ctx->storedJ_rows = (size_t)R;
  ctx->storedJ_cols = (size_t)N;
  return hipSuccess;
}

hipError_t launch_storedj_apply(jgm_context* ctx, int gemm_precision, int transpose,
                                const float* in, size_t nrhs, float* out) {
  const long R = (long)ctx->storedJ_rows, N = (long)ctx->storedJ_cols;
  const int P = (int)((nrhs + 63) / 64 * 64);
  const long K = transpose ? R : N;  // length of each right-hand side
  hipStream_t st = ctx->stream;
  hipError_t e = ctx->coef32.ensure(sizeof(float) * (size_t)K * P);
  if (e != hipSuccess) return e;
  float* Xt = ctx->coef32.as<float>();
  hipLaunchKernelGGL(k_transpose_pad, dim3(blocks_for(K * P)), dim3(kBlock), 0, st, in, K,
                     (int)nrhs, P, Xt);
  const float* J = ctx->storedJ.as<const float>();
  if (!transpose) {
    const long row_tiles = (R + 127) / 128;
    const int rhs_tiles = P / 64;
    long nsplit = std::max<long>(1, 2048 / std::max<long>(1, row_tiles * rhs_tiles));
    nsplit = std::min<long>(nsplit, std::max<long>(1, N / 256));
    long per = (N + nsplit - 1) / nsplit;
    per = (per + 15) / 16 * 16;  // whole K steps of every precision
    nsplit = (N + per - 1) / per;
    if ((e = ctx->partial.ensure(sizeof(float) * (size_t)nsplit * R * P)) != hipSuccess) return e;
    dim3 grid((unsigned)row_tiles, (unsigned)nsplit, (unsigned)rhs_tiles);
    {
      ProfScope ps(ctx, gemm_precision == JGM_GEMM_FP32   ? "storedj_jm_fp32"
                        : gemm_precision == JGM_GEMM_BF16 ? "storedj_jm_bf16"
                                                          : "storedj_jm_bf16_refined");
      switch (gemm_precision) {
        case JGM_GEMM_FP32:
          hipLaunchKernelGGL((k_gemm_jm<JGM_GEMM_FP32>), grid, dim3(kBlock), 0, st, J, R, N, Xt,
                             P, per, ctx->partial.as<float>());
          break;
        case JGM_GEMM_BF16:
          hipLaunchKernelGGL((k_gemm_jm<JGM_GEMM_BF16>), grid, dim3(kBlock), 0, st, J, R, N, Xt,
                             P, per, ctx->partial.as<float>());
          break;
        default:
          hipLaunchKernelGGL((k_gemm_jm<JGM_GEMM_BF16_REFINED>), grid, dim3(kBlock), 0, st, J, R,
                             N, Xt, P, per, ctx->partial.as<float>());
          break;
      }
    }
    hipLaunchKernelGGL(k_gemm_jm_reduce, dim3(blocks_for(R * (long)nrhs)), dim3(kBlock), 0, st,
                       ctx->partial.as<const float>(), (int)nsplit, R, P, (int)nrhs, out);
  } else {
    dim3 grid((unsigned)((N + 127) / 128), (unsigned)(P / 64));
    ProfScope ps(ctx, gemm_precision == JGM_GEMM_FP32   ? "storedj_jtr_fp32"
                      : gemm_precision == JGM_GEMM_BF16 ? "storedj_jtr_bf16"
                                                        : "storedj_jtr_bf16_refined");
    switch (gemm_precision) {
      case JGM_GEMM_FP32:
        hipLaunchKernelGGL((k_gemm_jtr<JGM_GEMM_FP32>), grid, dim3(kBlock), 0, st, J, R, N, Xt, P,
                           out, (int)nrhs);
        break;
      case JGM_GEMM_BF16:
        hipLaunchKernelGGL((k_gemm_jtr<JGM_GEMM_BF16>), grid, dim3(kBlock), 0, st, J, R, N, Xt, P,
                           out, (int)nrhs);
        break;
      default:
        hipLaunchKernelGGL((k_gemm_jtr<JGM_GEMM_BF16_REFINED>), grid, dim3(kBlock), 0, st, J, R,
                           N, Xt, P, out, (int)nrhs);
        break;
    }
  }
  return hipGetLastError();
}

}  // namespace jgm
