// Stored-sensitivity path: J built once in HBM (FP32), then applied to many
// right-hand sides at once on the matrix cores.
//
// J layout: column-major, J[c * ld + row], rows = site * kJ + component fastest, with
// the compact components per site: gz (scalar); Uxx, Uxy, Uxz, Uyy, Uyz, Uzz
// (tensor); Bx, By, Bz (magnetic); gz then the six tensor values (scalar+tensor).
// Columns are the mesh cells (background layers are not stored).  The leading
// dimension ld is the row count rounded up to 64 and the column count is padded to a
// multiple of 256; padding entries are zero, so the product kernels read whole tiles
// without bounds tests.
//
// Products (nrhs right-hand sides, column-major operands):
//   Y = J M    (R x nrhs):  split over cells.  A wave owns 64 rows x 64 right-hand
//                          sides (2 x 2 MFMA tiles); its A fragments are 128-byte row
//                          segments of one J column, read straight from HBM with the
//                          next 16-cell chunk in flight while the current one is
//                          multiplied.  Partial sums per split are reduced in FP64.
//   G = J^T Rr (N x nrhs):  a workgroup owns 256 cells x 64 right-hand sides and walks
//                          all rows in stages of 32; each stage of J (32 KB, one
//                          128-byte line per cell) and of Rr passes through LDS, the
//                          next stage's global loads overlapping this stage's MFMAs.
// M and Rr are first transposed to [k][nrhs_padded] (FP32, nrhs padded to 64, k padded
// with zero rows to the padded J extent).
// Precisions: FP32 operands on v_mfma_f32_32x32x2_f32; BF16 operands on
// v_mfma_f32_32x32x16_bf16 (FP32 accumulate), with the "refined" variant splitting
// both operands into hi + lo BF16 parts in registers and summing hi*hi + hi*lo +
// lo*hi (per-product error <= 3 * 2^-16 relative instead of 2^-7).
#include <type_traits>
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <string>
#include <vector>

#include "context.hpp"
#include "epilogue.hpp"
#include "prism_device.hpp"

namespace jgm {

namespace {

constexpr int kBlock = 256;
constexpr int kRowPad = 64;    // J leading dimension multiple
constexpr int kColPad = 256;   // J column count multiple (one J^T R workgroup)
constexpr int kChunk = 16;     // cells per pipeline chunk of J M
constexpr int kStage = 32;     // rows per LDS stage of J^T R
constexpr int kPitch = 257;    // LDS row pitch (floats) of the J^T R stage
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline long round_up(long v, long m) { return (v + m - 1) / m * m; }

__host__ __device__ inline int compact_rows(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return 1;
    case JGM_KIND_TENSOR: return 6;
    case JGM_KIND_MAGNETIC: return 3;
    case JGM_KIND_MAGNETIC_GRADIENT: return 1;
    case JGM_KIND_MAGNETIZATION: return 3;
    default: return 7;
  }
}

// parameter columns of the stored J: three per cell for the magnetization kind
inline long storedj_columns(const jgm_context* ctx, int kind) {
  const long n = (long)ncells(ctx);
  return kind == JGM_KIND_MAGNETIZATION ? 3 * n : n;
}

// J streamed block by block from host memory or a file (not resident in HBM)
inline bool storedj_streamed(const jgm_context* ctx) {
  return ctx->storedJ_host != nullptr || ctx->storedJ_fd >= 0;
}

// the vertical-gradient and magnetization kinds have no background columns
__host__ __device__ inline bool has_background(int kind) {
  return kind != JGM_KIND_MAGNETIC_GRADIENT && kind != JGM_KIND_MAGNETIZATION;
}

int mask_of(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return kMaskScalar;
    case JGM_KIND_SCALAR_TENSOR: return kMaskAll;
    default: return kMaskTensor;
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_zero_f32(float* __restrict__ p, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// Build: one lane per site, the wave walks a chunk of columns (wave-uniform); FP64
// closed form per cell (8 corners), stored as T (FP32, or FP64 for the FP64 J) in the
// compact layout.  Column col
// is cell col, except for the magnetization kind, where it is component p = col / N of
// cell col % N (J[3 s + i][p N + c] = mscale T_ip, OMPMagnetizationImp.cpp:25-117); the
// vertical-gradient kind stores Bz at the site minus Bz 1 m below
// (OMPMagneticGradImp.cpp:13-106), mscale = F 1e-7 / G.
template <int MASK, class T>
__global__ __launch_bounds__(kBlock) void k_storedj_build(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const double* __restrict__ xlo, const double* __restrict__ ylo,
    const double* __restrict__ zlo, const double* __restrict__ xsz,
    const double* __restrict__ ysz, const double* __restrict__ zsz, int ny, int nz, long ncell,
    long cbase, long cend, long cells_per_chunk, int kind, FieldDir fd, double mscale, long ld,
    T* __restrict__ J) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = s < S ? s : S - 1;
  const double px = sx[sl], py = sy[sl], pz = sz[sl];
  const int kJ = compact_rows(kind);
  // columns [cbase, cend); column col is stored at J + (col - cbase) ld
  const long c0 = cbase + (long)blockIdx.y * cells_per_chunk;
  const long c1 = c0 + cells_per_chunk < cend ? c0 + cells_per_chunk : cend;
  for (long col = c0; col < c1; ++col) {
    const int p = kind == JGM_KIND_MAGNETIZATION ? (int)(col / ncell) : 0;
    const long c = col - p * ncell;
    const int k = (int)(c % nz);
    const long t = c / nz;
    const int j = (int)(t % ny);
    const int i = (int)(t / ny);
    double v[NC], u[7];
    box_terms_f64<MASK>(px, py, pz, xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
    spread_components<MASK>(v, u);  // gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz
    double w[7];
    if (MASK == kMaskTensor && kind == JGM_KIND_MAGNETIC_GRADIENT) {
      box_terms_f64<MASK>(px, py, pz + 1.0, xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
      spread_components<MASK>(v, w);
    }
    if (s >= S) continue;
    T* o = J + (col - cbase) * ld + (long)s * kJ;
    if (kind == JGM_KIND_MAGNETIC_GRADIENT) {
      const double b0 = (u[5] * fd.bx + u[6] * fd.by + u[3] * fd.bz) * mscale;
      const double b1 = (w[5] * fd.bx + w[6] * fd.by + w[3] * fd.bz) * mscale;
      o[0] = (T)(b0 - b1);
    } else if (kind == JGM_KIND_MAGNETIZATION) {
      // column p of T = [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
      const double t0 = p == 0 ? u[1] : p == 1 ? u[4] : u[5];
      const double t1 = p == 0 ? u[4] : p == 1 ? u[2] : u[6];
      const double t2 = p == 0 ? u[5] : p == 1 ? u[6] : u[3];
      o[0] = (T)(t0 * mscale);
      o[1] = (T)(t1 * mscale);
      o[2] = (T)(t2 * mscale);
    } else if (kind == JGM_KIND_SCALAR) {
      o[0] = (T)u[0];
    } else if (kind == JGM_KIND_MAGNETIC) {
      o[0] = (T)((u[1] * fd.bx + u[4] * fd.by + u[5] * fd.bz) * fd.scale);
      o[1] = (T)((u[4] * fd.bx + u[2] * fd.by + u[6] * fd.bz) * fd.scale);
      o[2] = (T)((u[5] * fd.bx + u[6] * fd.by + u[3] * fd.bz) * fd.scale);
    } else {
      int q = 0;
      if (kind == JGM_KIND_SCALAR_TENSOR) o[q++] = (T)u[0];
      o[q + 0] = (T)u[1];
      o[q + 1] = (T)u[4];
      o[q + 2] = (T)u[5];
      o[q + 3] = (T)u[2];
      o[q + 4] = (T)u[6];
      o[q + 5] = (T)u[3];
    }
  }
}

// Column-major X (n x nrhs) -> Xt[k][P] for k < npad (P = nrhs padded to 64); rows
// k >= n and columns b >= nrhs are zero.
template <class T>
__global__ __launch_bounds__(kBlock) void k_transpose_pad(const T* __restrict__ X, long n,
                                                          long npad, int nrhs, int P,
                                                          T* __restrict__ Xt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npad * P) return;
  const long k = idx / P;
  const int b = (int)(idx % P);
  Xt[idx] = (k < n && b < nrhs) ? X[k + (long)b * n] : T(0);
}

__device__ __forceinline__ f32x16 mfma32(float a, float b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// x = hi + lo + e, |e| <= 2^-16 |x|: hi is x rounded to BF16, lo the rounded remainder
// (x - hi is exact in FP32)
template <bool LO>
__device__ __forceinline__ void to_bf16(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const __bf16 h = (__bf16)x[e];
    hi[e] = h;
    if constexpr (LO) lo[e] = (__bf16)(x[e] - (float)h);
  }
}

// acc[ta][tb] += A_ta B_tb for a 2 x 2 block of 32x32 tiles.  Element s of a lane's
// operands is k-step s (FP32: the lane's k is lane/32 within each step) or, for BF16,
// element s of the single 16-deep step (k = 8 * (lane/32) + s); A and B elements with
// the same index always refer to the same k.
template <int PREC>
__device__ __forceinline__ void mma_block(const float (&a)[2][8], const float (&b)[2][8],
                                          f32x16 (&acc)[2][2]) {
  if constexpr (PREC == JGM_GEMM_FP32) {
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) acc[ta][tb] = mfma32(a[ta][s], b[tb][s], acc[ta][tb]);
  } else {
    constexpr bool R = PREC == JGM_GEMM_BF16_REFINED;
    bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      to_bf16<R>(a[t], ah[t], al[t]);
      to_bf16<R>(b[t], bh[t], bl[t]);
    }
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) {
        acc[ta][tb] = mfma_bf16(ah[ta], bh[tb], acc[ta][tb]);
        if constexpr (R) {
          acc[ta][tb] = mfma_bf16(ah[ta], bl[tb], acc[ta][tb]);
          acc[ta][tb] = mfma_bf16(al[ta], bh[tb], acc[ta][tb]);
        }
      }
  }
}

// Two-level accumulation: the MFMA accumulators are added into FP32 totals and reset
// every 256 k, so no FP32 running sum grows longer than 256 k (the error of one long
// FP32 sum grows with its length).
__device__ __forceinline__ void flush(f32x16 (&acc)[2][2], float (&tot)[2][2][16]) {
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        tot[ta][tb][r] += acc[ta][tb][r];
        acc[ta][tb][r] = 0.0f;
      }
}

// k index (within a 16-deep chunk) of operand element s for a lane in half lk
template <int PREC>
__device__ __forceinline__ int chunk_k(int s, int lk) {
  return PREC == JGM_GEMM_FP32 ? 2 * s + lk : 8 * lk + s;
}

// Workgroup w of nw runs on XCD w % 8: each XCD takes a contiguous range of the (split,
// row tile) items in split-major order, so that the row tiles of a split, which read the
// same cells of M, share one L2 (XCD x holds nw / 8 items, one more when x < nw % 8, which
// is the number of w < nw with w % 8 == x, so the map is one to one).
__device__ __forceinline__ void xcd_item(int w, int nw, int row_tiles, int& split, int& rt) {
  const int x = w & 7;
  const int item = x * (nw / 8) + min(x, nw % 8) + (w >> 3);
  split = item / row_tiles;
  rt = item % row_tiles;
}

// Y partials: Yp[split][row][P].  Grid: x = nsplit * row_tiles workgroups in XCD order
// (xcd_item), y = right-hand-side tiles of 64.  Each split covers `per` cells (a
// multiple of 64).
struct JmFrag {
  float a[2][8], b[2][8];
};
// register stages of k_gemm_jm: 4 (234 VGPRs for FP32, no spills) reads FP32 J at 113
// TFLOP/s on 7168 x 131072 with the clocks up; 2 reaches 84, 3 97; 5 and 6 need AGPR
// copies and are slower
constexpr int kJmStages = 4;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// The BF16-refined product may take M already split (MPRE: its high and low BF16 parts Mh,
// Ml in the [k / 8][P] layout of k_rows_to_bf16_hilo, two 16-byte loads per tile and chunk
// instead of sixteen 4-byte loads and their conversions); the same values and MFMA order.
struct JmFragPre {
  float a[2][8];
  bf16x8 bh[2], bl[2];
};

template <int PREC, bool MPRE = false>
__global__ __launch_bounds__(kBlock) void k_gemm_jm(const float* __restrict__ J, long ld, long R,
                                                    long colpad, const float* __restrict__ Mt,
                                                    const bf16x8* __restrict__ Mh,
                                                    const bf16x8* __restrict__ Ml, int P, long per,
                                                    int row_tiles, int nsplit,
                                                    float* __restrict__ Yp) {
  static_assert(!MPRE || PREC == JGM_GEMM_BF16_REFINED, "split M: the refined product only");
  const int w = blockIdx.x;
  int split, rt;
  xcd_item(w, nsplit * row_tiles, row_tiles, split, rt);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int li = lane & 31, lk = lane >> 5;
  const long row0 = (long)rt * 256 + wave * 64;
  if (row0 >= ld) return;  // whole wave outside the padded rows; no barriers follow
  const int b0 = blockIdx.y * 64;
  const long c0 = (long)split * per;
  const long cend = c0 + per < colpad ? c0 + per : colpad;
  const int nchunk = cend > c0 ? (int)((cend - c0) / kChunk) : 0;
  const float* Jw = J + row0 + li;
  const float* Mw = Mt + b0 + li;
  f32x16 acc[2][2];
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[ta][tb][r] = 0.0f;
  float tot[2][2][16];
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 16; ++r) tot[ta][tb][r] = 0.0f;
  // a ring of kJmStages chunks in registers: chunk q is loaded while q - kJmStages + 1 is
  // multiplied, so kJmStages - 1 chunks of loads are in flight behind each MFMA block
  using Frag = std::conditional_t<MPRE, JmFragPre, JmFrag>;
  Frag f[kJmStages];
  auto load = [&](Frag& g, int q) {
    const long cb = c0 + (long)(q < nchunk ? q : nchunk - 1) * kChunk;  // past the end: unused
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const long cell = cb + chunk_k<PREC>(s, lk);
      g.a[0][s] = __builtin_nontemporal_load(Jw + cell * ld);
      g.a[1][s] = __builtin_nontemporal_load(Jw + cell * ld + 32);
      if constexpr (!MPRE) {
        g.b[0][s] = Mw[cell * P];
        g.b[1][s] = Mw[cell * P + 32];
      }
    }
    if constexpr (MPRE) {
      // the lane's eight cells cb + 8 lk .. cb + 8 lk + 7 (k = 8 lk + s) of right-hand
      // sides b0 + li and b0 + li + 32: one group of the [k / 8][P] copies each
      const long grp = (cb / 8 + lk) * P + b0 + li;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        g.bh[t] = Mh[grp + 32 * t];
        g.bl[t] = Ml[grp + 32 * t];
      }
    }
  };
  // sched_barrier keeps the scheduler from sinking the loads ahead down to their uses,
  // which would expose the full memory latency once per chunk
  if (nchunk > 0) {
#pragma unroll
    for (int d = 0; d < kJmStages - 1; ++d) load(f[d], d);
    for (int q = 0; q < nchunk; q += kJmStages) {
#pragma unroll
      for (int d = 0; d < kJmStages; ++d) {
        load(f[(d + kJmStages - 1) % kJmStages], q + d + kJmStages - 1);
        __builtin_amdgcn_sched_barrier(0);
        if (q + d < nchunk) {
          if constexpr (MPRE) {
            // mma_block<BF16_REFINED> with M's parts loaded instead of formed
            bf16x8 ah[2], al[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) to_bf16<true>(f[d].a[t], ah[t], al[t]);
#pragma unroll
            for (int ta = 0; ta < 2; ++ta)
#pragma unroll
              for (int tb = 0; tb < 2; ++tb) {
                acc[ta][tb] = mfma_bf16(ah[ta], f[d].bh[tb], acc[ta][tb]);
                acc[ta][tb] = mfma_bf16(ah[ta], f[d].bl[tb], acc[ta][tb]);
                acc[ta][tb] = mfma_bf16(al[ta], f[d].bh[tb], acc[ta][tb]);
              }
          } else {
            mma_block<PREC>(f[d].a, f[d].b, acc);
          }
          if (((q + d + 1) & 15) == 0) flush(acc, tot);  // every 256 cells
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  flush(acc, tot);
  // D layout of a 32x32 MFMA: register r of lane l holds (row 8*(r/4) + 4*(l/32) + r%4, col l%32)
  float* out = Yp + (long)split * R * P;
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long row = row0 + 32 * ta + 8 * (r >> 2) + 4 * lk + (r & 3);
      if (row < R) {
        out[row * P + b0 + li] = tot[ta][0][r];
        out[row * P + b0 + 32 + li] = tot[ta][1][r];
      }
    }
}

// Operand blocks shared by the four waves of a workgroup: 64 k (eight groups of eight) x 64
// right-hand sides of a BF16 operand in the [k / 8][P] layout of k_rows_to_bf16, staged in
// LDS as [group][right-hand side].  Thread t carries entries t and t + 256 from global memory
// to LDS; a lane's MFMA operand for k-step st and tile t is entry (2 st + lk) 64 + 32 t + li.
struct OpStage {
  bf16x8 v[2];
};

__device__ __forceinline__ void op_load(OpStage& o, const bf16x8* __restrict__ base, long g0,
                                        int P) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = (int)threadIdx.x + 256 * h;
    o.v[h] = base[(g0 + (e >> 6)) * P + (e & 63)];
  }
}

__device__ __forceinline__ void op_store(const OpStage& o, bf16x8* lds) {
  lds[threadIdx.x] = o.v[0];
  lds[threadIdx.x + 256] = o.v[1];
}

// BF16 J M from the tiled BF16 copy of J (k_pack_bf16_rowtiles: per 256-row tile and 64-cell
// block the 2048 16-byte groups in the order the four waves' MFMA operands take them) and a
// BF16 copy of M (k_rows_to_bf16: eight cells of a right-hand side per 16 bytes).  Per
// 16-cell k-step a lane's A operand (row li or li + 32 of its 64, cells 8 lk .. 8 lk + 7) and
// B operand (right-hand side li or li + 32, the same cells) are one 16-byte load each, and
// each load instruction of a wave reads one contiguous 1 KB run (a workgroup's whole share
// of J is one contiguous range); J is loaded in 64-cell blocks (four k-steps) three blocks
// ahead of the multiplication, M two.  Same split layout, k order, rounding and flush
// interval as k_gemm_jm<BF16>, so the same results; per must be a multiple of 64.
__global__ __launch_bounds__(kBlock) void k_gemm_jm_bf16p(const bf16x8* __restrict__ Jt, long ld,
                                                          long R, long colpad,
                                                          const bf16x8* __restrict__ Mb, int P,
                                                          long per, int row_tiles, int nsplit,
                                                          float* __restrict__ Yp) {
  __shared__ bf16x8 sM[2][512];
  int split, rt;
  xcd_item(blockIdx.x, nsplit * row_tiles, row_tiles, split, rt);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int li = lane & 31, lk = lane >> 5;
  const long row0 = (long)rt * 256 + wave * 64;
  // a wave past the padded rows only helps stage M (every wave reaches every barrier)
  const bool active = row0 < ld;
  const int b0 = blockIdx.y * 64;
  const long c0 = (long)split * per;
  const long cend = c0 + per < colpad ? c0 + per : colpad;
  const long nblk = cend > c0 ? (cend - c0) / 64 : 0;
  // this wave's groups of its row tile's first block of the split (tiled_group_index)
  const bf16x8* A0 = Jt + ((long)rt * (colpad / 64) + c0 / 64) * 2048 + wave * 512 + lane;
  const bf16x8* Mbase = Mb + b0;
  f32x16 acc[2][2];
  float tot[2][2][16];
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[ta][tb][r] = tot[ta][tb][r] = 0.0f;
  // J fragments of four 64-cell blocks are live (loaded three blocks ahead); M is staged in
  // LDS for the four waves, block b + 1 stored while block b is multiplied, block b + 2 in
  // registers on its way
  struct JQ {
    bf16x8 a[4][2];  // [k-step][row tile]
  } j0, j1, j2, j3;
  OpStage m0, m1;
  auto loadJ = [&](JQ& q, long blk) {
    if (!active) return;
    const bf16x8* p = A0 + blk * 2048;
#pragma unroll
    for (int st = 0; st < 4; ++st)
#pragma unroll
      for (int ta = 0; ta < 2; ++ta)
        q.a[st][ta] = __builtin_bit_cast(
            bf16x8, __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p + st * 128 + ta * 64)));
  };
  auto clampb = [&](long b) { return b < nblk ? b : nblk - 1; };
  // block b (J set jc, M in LDS buffer buf): load J block b + 3 into jl and M block b + 2
  // into mload, multiply, store M block b + 1 (mstore) into the other buffer
  auto step = [&](const JQ& jc, JQ& jl, OpStage& mload, const OpStage& mstore, int buf, long b) {
    loadJ(jl, clampb(b + 3));
    op_load(mload, Mbase, (c0 + 64 * clampb(b + 2)) / 8, P);
    __builtin_amdgcn_sched_barrier(0);
    if (b < nblk && active) {
      const bf16x8* sb = sM[buf];
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const bf16x8 bv0 = sb[(2 * st + lk) * 64 + li], bv1 = sb[(2 * st + lk) * 64 + 32 + li];
#pragma unroll
        for (int ta = 0; ta < 2; ++ta) {
          acc[ta][0] = mfma_bf16(jc.a[st][ta], bv0, acc[ta][0]);
          acc[ta][1] = mfma_bf16(jc.a[st][ta], bv1, acc[ta][1]);
        }
      }
      if ((b & 3) == 3) flush(acc, tot);  // every 256 cells from c0
    }
    op_store(mstore, sM[buf ^ 1]);
    __syncthreads();
  };
  if (nblk > 0) {
    loadJ(j0, 0);
    loadJ(j1, clampb(1));
    loadJ(j2, clampb(2));
    op_load(m0, Mbase, c0 / 8, P);
    op_store(m0, sM[0]);
    op_load(m1, Mbase, (c0 + 64 * clampb(1)) / 8, P);
    __syncthreads();
    // block b: J set b % 4, M register stage b % 2, LDS buffer b % 2
    for (long q = 0; q < nblk; q += 4) {
      step(j0, j3, m0, m1, 0, q);
      step(j1, j0, m1, m0, 1, q + 1);
      step(j2, j1, m0, m1, 0, q + 2);
      step(j3, j2, m1, m0, 1, q + 3);
    }
  }
  flush(acc, tot);
  if (!active) return;
  float* out = Yp + (long)split * R * P;
#pragma unroll
  for (int ta = 0; ta < 2; ++ta)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const long row = row0 + 32 * ta + 8 * (r >> 2) + 4 * lk + (r & 3);
      if (row < R) {
        out[row * P + b0 + li] = tot[ta][0][r];
        out[row * P + b0 + 32 + li] = tot[ta][1][r];
      }
    }
}

// BF16 copy of J for the BF16 J M, tiled for the matrix cores: rows in tiles of 256 (a J M
// workgroup), cells in blocks of 64 (a pipeline stage), and within a tile's block the 2048
// groups of eight cells of one row in the order [wave][k-step][row half][lane / 32][lane % 32]:
// row = 256 tile + 64 wave + 32 half + lane % 32, cells 64 block + 16 k-step + 8 (lane / 32)
// .. + 7.  Rows past ld are zero.  Rounded to nearest even as mma_block rounds J.
__global__ __launch_bounds__(kBlock) void k_pack_bf16_rowtiles(const float* __restrict__ J, long ld,
                                                               long colpad,
                                                               bf16x8* __restrict__ Jt) {
  const long nblk = colpad / 64, n = (ld + 255) / 256 * 256 * colpad / 8;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (long)gridDim.x * blockDim.x) {
    long t = idx;
    const int li = (int)(t & 31); t >>= 5;
    const int lk = (int)(t & 1); t >>= 1;
    const int ta = (int)(t & 1); t >>= 1;
    const int st = (int)(t & 3); t >>= 2;
    const int wave = (int)(t & 3); t >>= 2;
    const long blk = t % nblk, rt = t / nblk;
    const long row = rt * 256 + wave * 64 + ta * 32 + li;
    const long c0 = blk * 64 + st * 16 + lk * 8;
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (__bf16)(row < ld ? J[(c0 + e) * ld + row] : 0.0f);
    Jt[idx] = v;
  }
}

// BF16 copy of J for the BF16 J^T R, tiled for the matrix cores: eight rows of one cell per
// 16 bytes, and the 32 cells of a tile side by side for each 8-row group, so that group g
// of cell c is element ((c / 32) ngr + g) 32 + c % 32 (ngr = ld / 8) and one load
// instruction of a wave reads two contiguous 512-byte runs.  Rounded to nearest even as
// mma_block rounds J.
__global__ __launch_bounds__(kBlock) void k_pack_bf16_tiles(const float* __restrict__ J, long ld,
                                                            long colpad, bf16x8* __restrict__ Jt) {
  const long ngr = ld / 8, n = colpad * ngr;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (long)gridDim.x * blockDim.x) {
    const long li = idx & 31, t = idx >> 5;
    const long g = t % ngr, tile = t / ngr;
    const float* src = J + (tile * 32 + li) * ld + 8 * g;
    const f32x4 lo = *(const f32x4*)src, hi = *(const f32x4*)(src + 4);
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[e] = (__bf16)lo[e];
      v[4 + e] = (__bf16)hi[e];
    }
    Jt[idx] = v;
  }
}

// Rt[ld][P] (FP32) -> Rb[ld / 8][P] (BF16, eight rows of one right-hand side per 16 bytes)
__global__ __launch_bounds__(kBlock) void k_rows_to_bf16(const float* __restrict__ Rt, long ngroups,
                                                         int P, bf16x8* __restrict__ Rb) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ngroups * P) return;
  const long g = idx / P;
  const int b = (int)(idx - g * P);
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (__bf16)Rt[(8 * g + e) * P + b];
  Rb[idx] = v;
}

// k_rows_to_bf16 with the low parts too: Rh the rows rounded to BF16, Rl the rounded
// remainders (to_bf16<true>)
__global__ __launch_bounds__(kBlock) void k_rows_to_bf16_hilo(const float* __restrict__ Rt,
                                                              long ngroups, int P,
                                                              bf16x8* __restrict__ Rh,
                                                              bf16x8* __restrict__ Rl) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ngroups * P) return;
  const long g = idx / P;
  const int b = (int)(idx - g * P);
  float x[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) x[e] = Rt[(8 * g + e) * P + b];
  bf16x8 h, l;
  to_bf16<true>(x, h, l);
  Rh[idx] = h;
  Rl[idx] = l;
}

// BF16 G^T = Rt^T J from the tiled BF16 copy of J (k_pack_bf16_tiles): workgroup = 256 cells (wave: 64
// cells) x 64 right-hand sides, no LDS.  For each 16-row k-step a lane's eight rows of its
// cell (B operand, k = 8 (lane / 32) + e) and of its right-hand side (A operand) are one
// 16-byte load each; J is loaded in 64-row blocks (four k-steps) two blocks ahead of the
// multiplication, R one block ahead.  Same k order, rounding and flush interval as k_gemm_jtr<BF16>, so the
// same results.
__global__ __launch_bounds__(kBlock) void k_gemm_jtr_bf16c(const bf16x8* __restrict__ Jc, long ld,
                                                           long N, long nblk,
                                                           const bf16x8* __restrict__ Rb, int P,
                                                           float* __restrict__ G, long ldg,
                                                           int nrhs) {
  __shared__ bf16x8 sR[2][512];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 31, lk = lane >> 5;
  const int b0 = blockIdx.y * 64;
  const long ngr = ld / 8;           // 8-row groups per cell
  const long nb = ld / 64;           // 64-row blocks (ld is a multiple of 64)
  const bf16x8* Rbase = Rb + b0;
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long cbase = blk * kColPad + 64 * wave;
    const bf16x8* J0 = Jc + (cbase / 32) * ngr * 32 + lk * 32 + li;  // tile of cells cbase..
    const bf16x8* J1 = J0 + ngr * 32;                                // .. and cbase + 32..
    f32x16 acc[2][2];
    float tot[2][2][16];
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[ta][tb][r] = tot[ta][tb][r] = 0.0f;
    // J fragments of four 64-row blocks are live (loaded three blocks ahead); R is staged in
    // LDS for the four waves, block b + 1 stored while block b is multiplied, block b + 2
    // in registers on its way
    struct JQ {
      bf16x8 b[4][2];  // [k-step][cell tile]
    } j0, j1, j2, j3;
    OpStage r0, r1;
    auto loadJ = [&](JQ& q, long blk64) {
      const long g0 = blk64 * 8;  // first 8-row group of the block
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const long g = g0 + 2 * st;  // this lane's group: 16 st rows + 8 lk
        q.b[st][0] = __builtin_bit_cast(bf16x8, __builtin_nontemporal_load(
                                                    reinterpret_cast<const u32x4*>(J0 + g * 32)));
        q.b[st][1] = __builtin_bit_cast(bf16x8, __builtin_nontemporal_load(
                                                    reinterpret_cast<const u32x4*>(J1 + g * 32)));
      }
    };
    auto clampb = [&](long b) { return b < nb ? b : nb - 1; };
    auto step = [&](const JQ& jc, JQ& jl, OpStage& rload, const OpStage& rstore, int buf,
                    long b) {
      loadJ(jl, clampb(b + 3));  // past the end: unused reloads of valid memory
      op_load(rload, Rbase, 8 * clampb(b + 2), P);
      __builtin_amdgcn_sched_barrier(0);
      if (b < nb) {
        const bf16x8* sa = sR[buf];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const bf16x8 av0 = sa[(2 * st + lk) * 64 + li], av1 = sa[(2 * st + lk) * 64 + 32 + li];
#pragma unroll
          for (int tb = 0; tb < 2; ++tb) {
            acc[0][tb] = mfma_bf16(av0, jc.b[st][tb], acc[0][tb]);
            acc[1][tb] = mfma_bf16(av1, jc.b[st][tb], acc[1][tb]);
          }
        }
        if ((b & 3) == 3) flush(acc, tot);  // every 256 rows
      }
      op_store(rstore, sR[buf ^ 1]);
      __syncthreads();
    };
    loadJ(j0, 0);
    loadJ(j1, clampb(1));
    loadJ(j2, clampb(2));
    op_load(r0, Rbase, 0, P);
    op_store(r0, sR[0]);
    op_load(r1, Rbase, 8 * clampb(1), P);
    __syncthreads();
    // block b: J set b % 4, R register stage b % 2, LDS buffer b % 2
    for (long q = 0; q < nb; q += 4) {
      step(j0, j3, r0, r1, 0, q);
      step(j1, j0, r1, r0, 1, q + 1);
      step(j2, j1, r0, r1, 0, q + 2);
      step(j3, j2, r1, r0, 1, q + 3);
    }
    flush(acc, tot);
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) {
        const long c = cbase + 32 * tb + li;
        if (c < N) {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int b = b0 + 32 * ta + 8 * (r >> 2) + 4 * lk + (r & 3);
            if (b < nrhs) G[c + (long)b * ldg] = tot[ta][tb][r];
          }
        }
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

// G^T = Rt^T J over all rows: workgroup = 256 cells (wave: 64 cells) x 64 right-hand
// sides.  A operand = Rt stage (i = right-hand side, k = row), B operand = J stage
// (k = row, j = cell), both read from LDS.
// MPRE (the refined product only): R already split into its BF16 parts Rhi, Rlo in the
// [k / 8][P] layout of k_rows_to_bf16_hilo; a stage holds them in the R area of LDS as
// [group][rhs] (high parts, then low parts), a lane's A operand for a tile is one 16-byte
// read of each, and only J is split in registers; the same values and MFMA order.
template <int PREC, bool MPRE = false>
__global__ __launch_bounds__(kBlock) void k_gemm_jtr(const float* __restrict__ J, long ld, long N,
                                                     long nblk, const float* __restrict__ Rt,
                                                     const bf16x8* __restrict__ Rhi,
                                                     const bf16x8* __restrict__ Rlo, int P,
                                                     float* __restrict__ G, long ldg, int nrhs) {
  static_assert(!MPRE || PREC == JGM_GEMM_BF16_REFINED, "split R: the refined product only");
  static_assert(kStage * 64 * sizeof(float) == 2 * (kStage / 8) * 64 * sizeof(bf16x8),
                "the split R stage fills the R area");
  __shared__ float jt[kStage * kPitch];                           // [row][cell]
  __shared__ __attribute__((aligned(16))) float rt[kStage * 64];  // [row][rhs]
  bf16x8* const rth = reinterpret_cast<bf16x8*>(rt);              // MPRE: [group][rhs]
  bf16x8* const rtl = rth + (kStage / 8) * 64;
  const int b0 = blockIdx.y * 64;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int li = lane & 31, lk = lane >> 5;
  // loaders: J by 16-byte pieces (8 per cell line), Rt by 16-byte pieces (16 per row)
  const int jpart = tid & 7, jcell = tid >> 3;
  const int rpart = tid & 15, rrow = tid >> 4;
  const float* Rl = Rt + (long)rrow * P + b0 + rpart * 4;
  const long nst = ld / kStage;
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long cbase = blk * kColPad;
    const float* Jl = J + (cbase + jcell) * ld + jpart * 4;
    // two stages in flight (one workgroup per CU at these register counts): registers p0
    // and p1 hold the next two stages' pieces.  Measured against one stage on the same
    // GPU, J 7168 x 131072, 64 right-hand sides: FP32 1.23 -> 1.19 ms, BF16 0.72 -> 0.67 ms,
    // BF16-refined 0.83 -> 0.81 ms.
    f32x4 pj0[8], pr0[2], pj1[8], pr1[2];
    auto gload = [&](long st, f32x4 (&pj)[8], f32x4 (&pr)[2]) {
      const long r0 = (st < nst ? st : nst - 1) * kStage;  // past the end: unused reload
#pragma unroll
      for (int q = 0; q < 8; ++q)
        pj[q] = *(const f32x4*)(Jl + (long)q * 32 * ld + r0);
      if constexpr (MPRE) {
        // thread t: group t / 64 of the stage's four, right-hand side b0 + t % 64; the two
        // 16-byte parts travel in pr as raw bits
        const long gi = (r0 / 8 + (tid >> 6)) * P + b0 + (tid & 63);
        pr[0] = __builtin_bit_cast(f32x4, Rhi[gi]);
        pr[1] = __builtin_bit_cast(f32x4, Rlo[gi]);
      } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) pr[h] = *(const f32x4*)(Rl + (r0 + 16 * h) * P);
      }
    };
    f32x16 acc[2][2];  // [rhs tile][cell tile]
    float tot[2][2][16];
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[ta][tb][r] = tot[ta][tb][r] = 0.0f;
    // stage st: to LDS from (pj, pr), refill (pj, pr) with stage st + 2, multiply
    auto stage = [&](long st, f32x4 (&pj)[8], f32x4 (&pr)[2]) {
      __syncthreads();  // the previous stage has been consumed
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float* d = jt + jpart * 4 * kPitch + jcell + 32 * q;
        d[0] = pj[q][0];
        d[kPitch] = pj[q][1];
        d[2 * kPitch] = pj[q][2];
        d[3 * kPitch] = pj[q][3];
      }
      if constexpr (MPRE) {
        rth[tid] = __builtin_bit_cast(bf16x8, pr[0]);
        rtl[tid] = __builtin_bit_cast(bf16x8, pr[1]);
      } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) *(f32x4*)(rt + (rrow + 16 * h) * 64 + rpart * 4) = pr[h];
      }
      __syncthreads();
      gload(st + 2, pj, pr);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        float a[2][8], b[2][8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int row = 16 * half + chunk_k<PREC>(s, lk);
          if constexpr (!MPRE) {
            a[0][s] = rt[row * 64 + li];
            a[1][s] = rt[row * 64 + 32 + li];
          }
          b[0][s] = jt[row * kPitch + 64 * wave + li];
          b[1][s] = jt[row * kPitch + 64 * wave + 32 + li];
        }
        if constexpr (MPRE) {
          // mma_block<BF16_REFINED> with R's parts read instead of formed: the lane's rows
          // 16 half + 8 lk + s are group 2 half + lk
          const int g = 2 * half + lk;
          bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            ah[t] = rth[g * 64 + 32 * t + li];
            al[t] = rtl[g * 64 + 32 * t + li];
            to_bf16<true>(b[t], bh[t], bl[t]);
          }
#pragma unroll
          for (int ta = 0; ta < 2; ++ta)
#pragma unroll
            for (int tb = 0; tb < 2; ++tb) {
              acc[ta][tb] = mfma_bf16(ah[ta], bh[tb], acc[ta][tb]);
              acc[ta][tb] = mfma_bf16(ah[ta], bl[tb], acc[ta][tb]);
              acc[ta][tb] = mfma_bf16(al[ta], bh[tb], acc[ta][tb]);
            }
        } else {
          mma_block<PREC>(a, b, acc);
        }
      }
      if ((st & 7) == 7) flush(acc, tot);  // every 256 rows
    };
    gload(0, pj0, pr0);
    gload(1, pj1, pr1);
    for (long st = 0; st < nst; st += 2) {  // nst is even (ld is a multiple of 64)
      stage(st, pj0, pr0);
      stage(st + 1, pj1, pr1);
    }
    flush(acc, tot);
#pragma unroll
    for (int ta = 0; ta < 2; ++ta)
#pragma unroll
      for (int tb = 0; tb < 2; ++tb) {
        const long c = cbase + 64 * wave + 32 * tb + li;
        if (c < N) {
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int b = b0 + 32 * ta + 8 * (r >> 2) + 4 * lk + (r & 3);
            if (b < nrhs) G[c + (long)b * ldg] = tot[ta][tb][r];
          }
        }
      }
  }
}

// One model (or gradient) vector as the first right-hand side of a padded FP32
// operand Xt[npad][64]; the other 63 columns and the padding rows are zero.
// FP64 model vectors (model b at v + b * v_ld, b < nb) -> right-hand sides of
// Xt[npad][P] (FP32; columns nb.. and rows n.. zero).
template <class T>
__global__ __launch_bounds__(kBlock) void k_vec_to_rhs(const double* __restrict__ v, long v_ld,
                                                       int nb, long n, long npad, int P,
                                                       T* __restrict__ Xt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npad * P) return;
  const long k = idx / P;
  const int b = (int)(idx - k * P);
  Xt[idx] = (b < nb && k < n) ? (T)v[b * v_ld + k] : T(0);
}

// Raw misfits (raw layout per site; set b at rm + b * rm_ld, b < nb) -> compact
// residual rows as right-hand sides of Xt[ld][P]: the symmetric tensor entries appear
// twice in the raw layout, so their compact rows receive the sum of both residuals.
template <class T>
__global__ __launch_bounds__(kBlock) void k_rm_to_compact(const double* __restrict__ rm0,
                                                          long rm_ld, int nb, int S, int kind,
                                                          long ld, int P,
                                                          T* __restrict__ Xt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ld * P) return;
  const long row = idx / P;
  const int b = (int)(idx - row * P);
  const int kJ = compact_rows(kind);
  double v = 0.0;
  if (b < nb && row < (long)S * kJ) {
    const double* rm = rm0 + b * rm_ld;
    const long s = row / kJ;
    int q = (int)(row % kJ);
    if (kind == JGM_KIND_SCALAR || kind == JGM_KIND_MAGNETIC_GRADIENT) {
      v = rm[s];
    } else if (kind == JGM_KIND_MAGNETIC || kind == JGM_KIND_MAGNETIZATION) {
      v = rm[s * 3 + q];
    } else {
      const int kraw = kind == JGM_KIND_SCALAR_TENSOR ? 10 : 9;
      const double* r = rm + s * kraw;
      if (kind == JGM_KIND_SCALAR_TENSOR) {
        if (q == 0) {
          v = r[0];
          q = -1;
        } else {
          r += 1;
          q -= 1;
        }
      }
      switch (q) {  // compact Uxx, Uxy, Uxz, Uyy, Uyz, Uzz
        case 0: v = r[0]; break;
        case 1: v = (r[1] + r[3]); break;
        case 2: v = (r[2] + r[6]); break;
        case 3: v = r[4]; break;
        case 4: v = (r[5] + r[7]); break;
        case 5: v = r[8]; break;
        default: break;
      }
    }
  }
  Xt[idx] = (T)v;
}

// Cached forward epilogue: compact J m (FP32) of site s plus the background layers,
// assembled into the raw layout and transformed like the matrix-free path.
template <class T>
__global__ __launch_bounds__(kBlock) void k_storedj_finish(
    const T* __restrict__ Y, int S, int kind, const double* __restrict__ bgT,
    const double* __restrict__ bgvals, int L, FieldDir fd, int tf, double tfbx, double tfby,
    double tfbz, double* __restrict__ out, double* __restrict__ raw_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int kJ = compact_rows(kind);
  const T* y = Y + (long)s * kJ;
  if (!has_background(kind)) {  // J rows are the data themselves
    double raw[3];
    for (int m = 0; m < kJ; ++m) raw[m] = y[m];
    store_site(raw, kJ, s, tf, tfbx, tfby, tfbz, out, raw_out);
    return;
  }
  double u[7] = {0, 0, 0, 0, 0, 0, 0};  // gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz
  if (kind == JGM_KIND_SCALAR) {
    u[0] = y[0];
  } else if (kind == JGM_KIND_TENSOR || kind == JGM_KIND_SCALAR_TENSOR) {
    const int o = kind == JGM_KIND_SCALAR_TENSOR ? 1 : 0;
    if (o) u[0] = y[0];
    u[1] = y[o + 0];
    u[4] = y[o + 1];
    u[5] = y[o + 2];
    u[2] = y[o + 3];
    u[6] = y[o + 4];
    u[3] = y[o + 5];
  }
  double raw[10];
  const int kraw = assemble_raw(u, s, bgT, bgvals, L, kind, fd, raw);
  if (kind == JGM_KIND_MAGNETIC) {  // J rows are already the projected field
    raw[0] += y[0];
    raw[1] += y[1];
    raw[2] += y[2];
  }
  store_site(raw, kraw, s, tf, tfbx, tfby, tfbz, out, raw_out);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scale_to_f64(const T* __restrict__ g, long n,
                                                         double scale, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = scale * (double)g[i];
}

}  // namespace

namespace {

// Y += sum over splits of Yp[split][row][b] (FP64 totals of a J streamed in blocks)
__global__ __launch_bounds__(kBlock) void k_gemm_jm_accumulate(const float* __restrict__ Yp,
                                                               int nsplit, long R, int P,
                                                               int nrhs, double* __restrict__ Y) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= R * nrhs) return;
  const long row = idx / nrhs;
  const int b = (int)(idx % nrhs);
  double a = 0.0;
  for (int s = 0; s < nsplit; ++s) a += Yp[(long)s * R * P + row * P + b];
  Y[row + (long)b * R] += a;
}

__global__ __launch_bounds__(kBlock) void k_fill_zero_f64(double* __restrict__ p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(const double* __restrict__ a, long n,
                                                       float* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)a[i];
}

// Build the J columns of cells [c0, c1) into J (column c - c0 at J + (c - c0) * ld).
template <class T>
hipError_t build_columns(jgm_context* ctx, int kind, long c0, long c1, long ld, T* J) {
  const int S = (int)ctx->S;
  const int site_blocks = (S + kBlock - 1) / kBlock;
  const long n = c1 - c0;
  const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
  long want = std::max<long>(1, std::min<long>(n, (16384 + waves_per_chunk - 1) / waves_per_chunk));
  const long per = (n + want - 1) / want;
  const int nchunks = (int)((n + per - 1) / per);
  const int ny = (int)ctx->ny, nz = (int)ctx->nz;
  FieldDir fd = field_dir(ctx);
  // unit factors of the magnetic kinds that the J rows carry (the tensor terms include G)
  const double mu0 = 4e-7 * kPi;  // OMPMagnetizationImp.cpp:35
  const double mscale = kind == JGM_KIND_MAGNETIZATION ? 1.0e9 * mu0 / (4 * kPi * kGravConst)
                                                       : ctx->fieldstrength * (1e-7 / kGravConst);
  ProfScope ps(ctx, "storedj_build_fp64");
#define JGM_BUILD(M)                                                                            \
  hipLaunchKernelGGL((k_storedj_build<M, T>), dim3(site_blocks, nchunks), dim3(kBlock), 0,     \
                     ctx->stream, ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),               \
                     ctx->d_sz.as<double>(), S, ctx->d_xlo.as<double>(), ctx->d_ylo.as<double>(), \
                     ctx->d_zlo.as<double>(), ctx->d_xsz.as<double>(), ctx->d_ysz.as<double>(), \
                     ctx->d_zsz.as<double>(), ny, nz, (long)ncells(ctx), c0, c1, per, kind, fd, \
                     mscale, ld, J)
  switch (mask_of(kind)) {
    case kMaskScalar: JGM_BUILD(kMaskScalar); break;
    case kMaskAll: JGM_BUILD(kMaskAll); break;
    default: JGM_BUILD(kMaskTensor); break;
  }
#undef JGM_BUILD
  return hipGetLastError();
}

hipError_t ensure_streaming(jgm_context* ctx) {
  hipError_t e = hipSuccess;
  if (!ctx->copy_stream) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    if (!ctx->ev_copied[i]) e = hipEventCreateWithFlags(&ctx->ev_copied[i], hipEventDisableTiming);
    if (e == hipSuccess && !ctx->ev_used[i])
      e = hipEventCreateWithFlags(&ctx->ev_used[i], hipEventDisableTiming);
  }
  if (e == hipSuccess && !ctx->ev_start)
    e = hipEventCreateWithFlags(&ctx->ev_start, hipEventDisableTiming);
  return e;
}

// Calls fn(J_block, first cell, cells) for every column block of the stored J: once for
// a J in HBM; for a J in host memory, per block after its copy into one of two HBM
// staging buffers (copy stream), so that the copy of block b + 1 overlaps the products
// on block b (context stream).  A staging buffer is refilled only after the products
// that read it have finished.
// pread / pwrite of a whole byte range (short transfers repeated); false on an I/O error
bool file_io(int fd, void* buf, size_t bytes, off_t off, bool write) {
  char* p = static_cast<char*>(buf);
  while (bytes > 0) {
    const ssize_t n = write ? ::pwrite(fd, p, bytes, off) : ::pread(fd, p, bytes, off);
    if (n < 0 && errno == EINTR) continue;
    if (n <= 0) return false;
    p += n;
    bytes -= (size_t)n;
    off += n;
  }
  return true;
}

// The file-backed J: block b read from the file into one of two page-locked buffers
// while the GPU works on block b - 1; each buffer is reread only after its copy to HBM
// has completed.
template <class F>
hipError_t for_each_file_block(jgm_context* ctx, F&& fn) {
  const long ld = (long)ctx->storedJ_ld, colpad = (long)ctx->storedJ_colpad;
  const long blk = (long)ctx->storedJ_block;
  float* stage[2] = {ctx->storedJ.as<float>(), ctx->storedJ.as<float>() + blk * ld};
  hipError_t e;
  if ((e = hipEventRecord(ctx->ev_start, ctx->stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_start, 0)) != hipSuccess) return e;
  auto read_block = [&](long c0, float* dst) -> hipError_t {
    const long nc = std::min(blk, colpad - c0);
    return file_io(ctx->storedJ_fd, dst, sizeof(float) * (size_t)(nc * ld),
                   (off_t)(sizeof(float) * (size_t)(c0 * ld)), false)
               ? hipSuccess
               : hipErrorFileNotFound;
  };
  if ((e = read_block(0, ctx->storedJ_pin[0])) != hipSuccess) return e;
  for (long c0 = 0, b = 0; c0 < colpad; c0 += blk, ++b) {
    const int i = (int)(b & 1);
    const long nc = std::min(blk, colpad - c0);
    if (b >= 2 && (e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_used[i], 0)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(stage[i], ctx->storedJ_pin[i], sizeof(float) * nc * ld,
                            hipMemcpyHostToDevice, ctx->copy_stream)) != hipSuccess)
      return e;
    if ((e = hipEventRecord(ctx->ev_copied[i], ctx->copy_stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(ctx->stream, ctx->ev_copied[i], 0)) != hipSuccess) return e;
    if ((e = fn(static_cast<const float*>(stage[i]), c0, nc)) != hipSuccess) return e;
    if ((e = hipEventRecord(ctx->ev_used[i], ctx->stream)) != hipSuccess) return e;
    if (c0 + blk < colpad) {  // the next block into the other buffer, once its copy is done
      if (b >= 1 && (e = hipEventSynchronize(ctx->ev_copied[1 - i])) != hipSuccess) return e;
      if ((e = read_block(c0 + blk, ctx->storedJ_pin[1 - i])) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

template <class F>
hipError_t for_each_block(jgm_context* ctx, F&& fn) {
  const long ld = (long)ctx->storedJ_ld, colpad = (long)ctx->storedJ_colpad;
  if (!storedj_streamed(ctx)) return fn(ctx->storedJ.as<const float>(), 0L, colpad);
  hipError_t e = ensure_streaming(ctx);
  if (e != hipSuccess) return e;
  if (ctx->storedJ_fd >= 0) return for_each_file_block(ctx, fn);
  const long blk = (long)ctx->storedJ_block;
  float* stage[2] = {ctx->storedJ.as<float>(), ctx->storedJ.as<float>() + blk * ld};
  // earlier work on the context stream may still read the staging buffers
  if ((e = hipEventRecord(ctx->ev_start, ctx->stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_start, 0)) != hipSuccess) return e;
  for (long c0 = 0, b = 0; c0 < colpad; c0 += blk, ++b) {
    const int i = (int)(b & 1);
    const long nc = std::min(blk, colpad - c0);
    if (b >= 2 && (e = hipStreamWaitEvent(ctx->copy_stream, ctx->ev_used[i], 0)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(stage[i], ctx->storedJ_host + c0 * ld, sizeof(float) * nc * ld,
                            hipMemcpyHostToDevice, ctx->copy_stream)) != hipSuccess)
      return e;
    if ((e = hipEventRecord(ctx->ev_copied[i], ctx->copy_stream)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(ctx->stream, ctx->ev_copied[i], 0)) != hipSuccess) return e;
    if ((e = fn(static_cast<const float*>(stage[i]), c0, nc)) != hipSuccess) return e;
    if ((e = hipEventRecord(ctx->ev_used[i], ctx->stream)) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

void release_host_storedj(jgm_context* ctx) {
  if (ctx->storedJ_host || ctx->storedJ_fd >= 0) {
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  }
  if (ctx->storedJ_host) (void)hipHostFree(ctx->storedJ_host);
  for (float*& p : ctx->storedJ_pin) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
  }
  if (ctx->storedJ_fd >= 0) {
    ::close(ctx->storedJ_fd);
    ::unlink(ctx->storedJ_file.c_str());
  }
  ctx->storedJ_fd = -1;
  ctx->storedJ_file.clear();
  ctx->storedJ_host = nullptr;
  ctx->storedJ_block = 0;
}

// J in a file in `dir`, built in blocks of block_cells cells through an HBM staging
// buffer and a page-locked buffer; the file is removed when J is released or rebuilt.
hipError_t launch_storedj_build_file(jgm_context* ctx, int kind, size_t block_cells,
                                     const char* dir) {
  static std::atomic<unsigned long> serial{0};
  const int S = (int)ctx->S;
  const long N = storedj_columns(ctx, kind);
  const long R = (long)S * compact_rows(kind);
  if (S == 0 || N == 0) return hipErrorInvalidValue;
  const long ld = round_up(R, kRowPad), colpad = round_up(N, kColPad);
  ctx->storedJ_kind = 0;
  ctx->storedJ_fp64 = false;
  ++ctx->storedJ_gen;
  ctx->storedJ_bf16.release();
  ctx->storedJ_bf16c.release();
  release_host_storedj(ctx);
  const long blk = std::min<long>(round_up((long)block_cells, kColPad), colpad);
  hipError_t e;
  ctx->storedJ.release();
  if ((e = ctx->storedJ.ensure(sizeof(float) * (size_t)(2 * blk * ld))) != hipSuccess) return e;
  for (float*& p : ctx->storedJ_pin) {
    void* h = nullptr;
    if ((e = hipHostMalloc(&h, sizeof(float) * (size_t)(blk * ld), hipHostMallocDefault)) !=
        hipSuccess) {
      release_host_storedj(ctx);
      return e;
    }
    p = static_cast<float*>(h);
  }
  ctx->storedJ_file = std::string(dir) + "/jgm_J_" + std::to_string((long)::getpid()) + "_" +
                      std::to_string(serial.fetch_add(1)) + ".bin";
  ctx->storedJ_fd = ::open(ctx->storedJ_file.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0600);
  if (ctx->storedJ_fd < 0) {
    const int err = errno;  // kept for the caller's message past the frees below
    ctx->storedJ_file.clear();
    release_host_storedj(ctx);
    errno = err;
    return hipErrorFileNotFound;
  }
  ctx->storedJ_block = (size_t)blk;
  float* stage = ctx->storedJ.as<float>();
  for (long c0 = 0; c0 < colpad && e == hipSuccess; c0 += blk) {
    const long nc = std::min(blk, colpad - c0);
    hipLaunchKernelGGL(k_fill_zero_f32, dim3(8 * ctx->cu_count), dim3(kBlock), 0, ctx->stream,
                       stage, nc * ld);
    if (c0 < N && (e = build_columns(ctx, kind, c0, std::min(c0 + nc, N), ld, stage)) !=
                      hipSuccess)
      break;
    if ((e = hipMemcpyAsync(ctx->storedJ_pin[0], stage, sizeof(float) * nc * ld,
                            hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
      break;
    if (!file_io(ctx->storedJ_fd, ctx->storedJ_pin[0], sizeof(float) * (size_t)(nc * ld),
                 (off_t)(sizeof(float) * (size_t)(c0 * ld)), true))
      e = hipErrorFileNotFound;
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    const int err = errno;
    release_host_storedj(ctx);
    errno = err;
    return e;
  }
  ctx->storedJ_kind = kind;
  ctx->storedJ_rows = (size_t)R;
  ctx->storedJ_cols = (size_t)N;
  ctx->storedJ_ld = (size_t)ld;
  ctx->storedJ_colpad = (size_t)colpad;
  return hipSuccess;
}

hipError_t launch_storedj_build(jgm_context* ctx, int kind, int precision) {
  return launch_storedj_build_resident(ctx, kind, precision, 0);
}

// block_cells == 0: J in HBM.  Otherwise J in page-locked host memory, built in blocks
// of block_cells cells (a multiple of 256) through an HBM staging buffer.
hipError_t launch_storedj_build_resident(jgm_context* ctx, int kind, int precision,
                                         size_t block_cells) {
  (void)precision;  // the build evaluates the FP64 closed form; J is stored in FP32
  const int S = (int)ctx->S;
  const long N = storedj_columns(ctx, kind);
  const int kJ = compact_rows(kind);
  const long R = (long)S * kJ;
  if (S == 0 || N == 0) return hipErrorInvalidValue;
  const long ld = round_up(R, kRowPad), colpad = round_up(N, kColPad);
  ctx->storedJ_kind = 0;
  ctx->storedJ_fp64 = false;
  ++ctx->storedJ_gen;  // the BF16 copies no longer match
  ctx->storedJ_bf16.release();
  ctx->storedJ_bf16c.release();
  release_host_storedj(ctx);
  hipError_t e;
  if (block_cells == 0) {
    if ((e = ctx->storedJ.ensure(sizeof(float) * (size_t)ld * colpad)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill_zero_f32, dim3(8 * ctx->cu_count), dim3(kBlock), 0, ctx->stream,
                       ctx->storedJ.as<float>(), ld * colpad);
    if ((e = build_columns(ctx, kind, 0, N, ld, ctx->storedJ.as<float>())) != hipSuccess)
      return e;
  } else {
    const long blk = std::min<long>(round_up((long)block_cells, kColPad), colpad);
    // the staging buffers replace the HBM copy of J
    ctx->storedJ.release();
    if ((e = ctx->storedJ.ensure(sizeof(float) * (size_t)(2 * blk * ld))) != hipSuccess) return e;
    void* host = nullptr;
    if ((e = hipHostMalloc(&host, sizeof(float) * (size_t)ld * colpad, hipHostMallocDefault)) !=
        hipSuccess)
      return e;
    ctx->storedJ_host = static_cast<float*>(host);
    ctx->storedJ_block = (size_t)blk;
    float* stage = ctx->storedJ.as<float>();
    for (long c0 = 0; c0 < colpad; c0 += blk) {
      const long nc = std::min(blk, colpad - c0);
      hipLaunchKernelGGL(k_fill_zero_f32, dim3(8 * ctx->cu_count), dim3(kBlock), 0, ctx->stream,
                         stage, nc * ld);
      if (c0 < N && (e = build_columns(ctx, kind, c0, std::min(c0 + nc, N), ld, stage)) !=
                        hipSuccess)
        break;
      if ((e = hipMemcpyAsync(ctx->storedJ_host + c0 * ld, stage, sizeof(float) * nc * ld,
                              hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess)
        break;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      release_host_storedj(ctx);
      return e;
    }
  }
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  ctx->storedJ_kind = kind;
  ctx->storedJ_rows = (size_t)R;
  ctx->storedJ_cols = (size_t)N;
  ctx->storedJ_ld = (size_t)ld;
  ctx->storedJ_colpad = (size_t)colpad;
  return hipSuccess;
}

// FP64 J in HBM (the precision of the reference's FullSensitivity matrix), same compact
// layout and padding as the FP32 J, for the FP64 matrix-core products.
hipError_t launch_storedj_build_f64(jgm_context* ctx, int kind) {
  const int S = (int)ctx->S;
  const long N = storedj_columns(ctx, kind);
  const long R = (long)S * compact_rows(kind);
  if (S == 0 || N == 0) return hipErrorInvalidValue;
  const long ld = round_up(R, kRowPad), colpad = round_up(N, kColPad);
  ctx->storedJ_kind = 0;
  ctx->storedJ_fp64 = false;
  ++ctx->storedJ_gen;
  ctx->storedJ_bf16.release();
  ctx->storedJ_bf16c.release();
  release_host_storedj(ctx);
  hipError_t e;
  if ((e = ctx->storedJ.ensure(sizeof(double) * (size_t)ld * colpad)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_zero_f64, dim3(8 * ctx->cu_count), dim3(kBlock), 0, ctx->stream,
                     ctx->storedJ.as<double>(), ld * colpad);
  if ((e = build_columns(ctx, kind, 0, N, ld, ctx->storedJ.as<double>())) != hipSuccess) return e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  ctx->storedJ_kind = kind;
  ctx->storedJ_fp64 = true;
  ctx->storedJ_rows = (size_t)R;
  ctx->storedJ_cols = (size_t)N;
  ctx->storedJ_ld = (size_t)ld;
  ctx->storedJ_colpad = (size_t)colpad;
  return hipSuccess;
}

namespace {

// Makes the packed BF16 copy of the HBM-resident J if it is missing or older than J.
// false when device memory for it (2 bytes per J element) cannot be had: the BF16
// product then rounds the FP32 J in registers, with the same results.
bool ensure_bf16_copy(jgm_context* ctx) {
  if (ctx->storedJ_bf16_gen == ctx->storedJ_gen) return true;
  if (!ctx->bf16_copy_enabled || ctx->storedJ_bf16_failed_gen == ctx->storedJ_gen) return false;
  const long ld = (long)ctx->storedJ_ld, colpad = (long)ctx->storedJ_colpad;
  // rows padded to whole 256-row tiles
  if (ctx->storedJ_bf16.ensure(sizeof(__bf16) * (size_t)round_up(ld, 256) * colpad) !=
      hipSuccess) {
    (void)hipGetLastError();
    ctx->storedJ_bf16.release();
    ctx->storedJ_bf16_failed_gen = ctx->storedJ_gen;  // not retried for this J
    return false;
  }
  hipLaunchKernelGGL(k_pack_bf16_rowtiles, dim3(8 * ctx->cu_count), dim3(kBlock), 0, ctx->stream,
                     ctx->storedJ.as<const float>(), ld, colpad, ctx->storedJ_bf16.as<bf16x8>());
  ctx->storedJ_bf16_gen = ctx->storedJ_gen;
  return true;
}

// The tiled BF16 copy for the BF16 J^T R, on the same terms.
bool ensure_bf16c_copy(jgm_context* ctx) {
  if (ctx->storedJ_bf16c_gen == ctx->storedJ_gen) return true;
  if (!ctx->bf16_copy_enabled || ctx->storedJ_bf16c_failed_gen == ctx->storedJ_gen) return false;
  const long ngroups = (long)ctx->storedJ_ld * (long)ctx->storedJ_colpad / 8;
  if (ctx->storedJ_bf16c.ensure(sizeof(bf16x8) * (size_t)ngroups) != hipSuccess) {
    (void)hipGetLastError();
    ctx->storedJ_bf16c.release();
    ctx->storedJ_bf16c_failed_gen = ctx->storedJ_gen;
    return false;
  }
  hipLaunchKernelGGL(k_pack_bf16_tiles, dim3(8 * ctx->cu_count), dim3(kBlock), 0, ctx->stream,
                     ctx->storedJ.as<const float>(), (long)ctx->storedJ_ld,
                     (long)ctx->storedJ_colpad, ctx->storedJ_bf16c.as<bf16x8>());
  ctx->storedJ_bf16c_gen = ctx->storedJ_gen;
  return true;
}

// J M on a transposed, padded operand Xt[colpad][P] -> out (R x nrhs, column-major)
hipError_t run_jm(jgm_context* ctx, int gemm_precision, const float* Xt, int P, int nrhs,
                  float* out) {
  const long R = (long)ctx->storedJ_rows;
  const long ld = (long)ctx->storedJ_ld;
  hipStream_t st = ctx->stream;
  const int rhs_tiles = P / 64;
  const int row_tiles = (int)((ld + 255) / 256);
  const bool streamed = storedj_streamed(ctx);
  hipError_t e;
  // the BF16 product of an HBM-resident J reads the packed BF16 copy, made once per J
  const bool packed = gemm_precision == JGM_GEMM_BF16 && !streamed && ensure_bf16_copy(ctx);
  if (streamed) {
    if ((e = ctx->acc64.ensure(sizeof(double) * (size_t)R * nrhs)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill_zero_f64, dim3(4 * ctx->cu_count), dim3(kBlock), 0, st,
                       ctx->acc64.as<double>(), R * (long)nrhs);
  }
  e = for_each_block(ctx, [&](const float* J, long c0, long ncols) -> hipError_t {
    // one workgroup per CU: a second co-resident workgroup slows both down (the CU's
    // memory pipeline is the limit)
    long nsplit = std::max<long>(1, (long)ctx->cu_count / ((long)row_tiles * rhs_tiles));
    nsplit = std::min<long>(nsplit, ncols / 32);
    // splits of whole 64-cell blocks (the packed kernel's stage)
    const long per = round_up((ncols + nsplit - 1) / nsplit, 4 * kChunk);
    hipError_t e2 = ctx->partial.ensure(sizeof(float) * (size_t)nsplit * R * P);
    if (e2 != hipSuccess) return e2;
    dim3 grid((unsigned)(nsplit * row_tiles), (unsigned)rhs_tiles);
    const float* Mt = Xt + c0 * P;
    if (packed) {  // M rounded to BF16 once, eight cells of a right-hand side per 16 bytes
      if ((e2 = ctx->storedJ_rhs16.ensure(sizeof(__bf16) * (size_t)ncols * P)) != hipSuccess)
        return e2;
      hipLaunchKernelGGL(k_rows_to_bf16, dim3(blocks_for(ncols / 8 * P)), dim3(kBlock), 0, st, Mt,
                         ncols / 8, P, ctx->storedJ_rhs16.as<bf16x8>());
    }
    // the refined product takes M split into its BF16 parts once per block instead of in
    // every workgroup that reads it
    const bool mpre = gemm_precision == JGM_GEMM_BF16_REFINED;
    if (mpre) {
      if ((e2 = ctx->storedJ_rhs16.ensure(sizeof(__bf16) * (size_t)ncols * P)) != hipSuccess ||
          (e2 = ctx->storedJ_rhs16lo.ensure(sizeof(__bf16) * (size_t)ncols * P)) != hipSuccess)
        return e2;
      hipLaunchKernelGGL(k_rows_to_bf16_hilo, dim3(blocks_for(ncols / 8 * P)), dim3(kBlock), 0, st,
                         Mt, ncols / 8, P, ctx->storedJ_rhs16.as<bf16x8>(),
                         ctx->storedJ_rhs16lo.as<bf16x8>());
    }
    {
      // the packed BF16 product is profiled under its own name, so that timings know
      // whether 2 or 4 bytes of J were read per element
      ProfScope ps(ctx, gemm_precision == JGM_GEMM_FP32   ? "storedj_jm_fp32"
                        : gemm_precision == JGM_GEMM_BF16 ? (packed ? "storedj_jm_bf16_packed"
                                                                    : "storedj_jm_bf16")
                                                          : "storedj_jm_bf16_refined");
#define JGM_JM(PR)                                                                        \
  hipLaunchKernelGGL((k_gemm_jm<PR>), grid, dim3(kBlock), 0, st, J, ld, R, ncols, Mt,         \
                     (const bf16x8*)nullptr, (const bf16x8*)nullptr, P, per, row_tiles,      \
                     (int)nsplit, ctx->partial.as<float>())
      switch (gemm_precision) {
        case JGM_GEMM_FP32: JGM_JM(JGM_GEMM_FP32); break;
        case JGM_GEMM_BF16:
          if (packed)
            hipLaunchKernelGGL(k_gemm_jm_bf16p, grid, dim3(kBlock), 0, st,
                               ctx->storedJ_bf16.as<const bf16x8>(), ld, R, ncols,
                               ctx->storedJ_rhs16.as<const bf16x8>(), P, per, row_tiles,
                               (int)nsplit, ctx->partial.as<float>());
          else
            JGM_JM(JGM_GEMM_BF16);
          break;
        default:
          hipLaunchKernelGGL((k_gemm_jm<JGM_GEMM_BF16_REFINED, true>), grid, dim3(kBlock), 0, st,
                             J, ld, R, ncols, Mt, ctx->storedJ_rhs16.as<const bf16x8>(),
                             ctx->storedJ_rhs16lo.as<const bf16x8>(), P, per, row_tiles,
                             (int)nsplit, ctx->partial.as<float>());
          break;
      }
#undef JGM_JM
    }
    if (streamed)
      hipLaunchKernelGGL(k_gemm_jm_accumulate, dim3(blocks_for(R * (long)nrhs)), dim3(kBlock), 0,
                         st, ctx->partial.as<const float>(), (int)nsplit, R, P, nrhs,
                         ctx->acc64.as<double>());
    else
      hipLaunchKernelGGL(k_gemm_jm_reduce, dim3(blocks_for(R * (long)nrhs)), dim3(kBlock), 0, st,
                         ctx->partial.as<const float>(), (int)nsplit, R, P, nrhs, out);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  if (streamed)
    hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(R * (long)nrhs)), dim3(kBlock), 0, st,
                       ctx->acc64.as<const double>(), R * (long)nrhs, out);
  return hipGetLastError();
}

// J^T R on a transposed, padded operand Xt[ld][P] -> out (N x nrhs, column-major)
hipError_t run_jtr(jgm_context* ctx, int gemm_precision, const float* Xt, int P, int nrhs,
                   float* out) {
  const long N = (long)ctx->storedJ_cols;
  const long ld = (long)ctx->storedJ_ld;
  hipStream_t st = ctx->stream;
  // the BF16 product of an HBM-resident J reads the column-major BF16 copy, made once per J
  if (gemm_precision == JGM_GEMM_BF16 && !storedj_streamed(ctx) && ensure_bf16c_copy(ctx)) {
    const long nblk = (long)ctx->storedJ_colpad / kColPad;
    hipError_t e = ctx->storedJ_rhs16.ensure(sizeof(__bf16) * (size_t)ld * P);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rows_to_bf16, dim3(blocks_for(ld / 8 * P)), dim3(kBlock), 0, st, Xt,
                       ld / 8, P, ctx->storedJ_rhs16.as<bf16x8>());
    ProfScope ps(ctx, "storedj_jtr_bf16_packed");
    hipLaunchKernelGGL(k_gemm_jtr_bf16c, dim3((unsigned)std::min<long>(nblk, 65535), (unsigned)(P / 64)),
                       dim3(kBlock), 0, st, ctx->storedJ_bf16c.as<const bf16x8>(), ld, N, nblk,
                       ctx->storedJ_rhs16.as<const bf16x8>(), P, out, N, nrhs);
    return hipGetLastError();
  }
  // the refined product takes R split into its BF16 parts once instead of in every
  // workgroup (R is the same for every block of J)
  if (gemm_precision == JGM_GEMM_BF16_REFINED) {
    hipError_t e;
    if ((e = ctx->storedJ_rhs16.ensure(sizeof(__bf16) * (size_t)ld * P)) != hipSuccess ||
        (e = ctx->storedJ_rhs16lo.ensure(sizeof(__bf16) * (size_t)ld * P)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_rows_to_bf16_hilo, dim3(blocks_for(ld / 8 * P)), dim3(kBlock), 0, st, Xt,
                       ld / 8, P, ctx->storedJ_rhs16.as<bf16x8>(),
                       ctx->storedJ_rhs16lo.as<bf16x8>());
  }
  return for_each_block(ctx, [&](const float* J, long c0, long ncols) -> hipError_t {
    const long nblk = ncols / kColPad;
    const long nvalid = std::max<long>(0, std::min(ncols, N - c0));
    if (nvalid == 0) return hipSuccess;
    dim3 grid((unsigned)std::min<long>(nblk, 65535), (unsigned)(P / 64));
    ProfScope ps(ctx, gemm_precision == JGM_GEMM_FP32   ? "storedj_jtr_fp32"
                      : gemm_precision == JGM_GEMM_BF16 ? "storedj_jtr_bf16"
                                                        : "storedj_jtr_bf16_refined");
#define JGM_JTR(PR, PRE)                                                                     \
  hipLaunchKernelGGL((k_gemm_jtr<PR, PRE>), grid, dim3(kBlock), 0, st, J, ld, nvalid, nblk, Xt, \
                     ctx->storedJ_rhs16.as<const bf16x8>(),                                 \
                     ctx->storedJ_rhs16lo.as<const bf16x8>(), P, out + c0, N, nrhs)
    switch (gemm_precision) {
      case JGM_GEMM_FP32: JGM_JTR(JGM_GEMM_FP32, false); break;
      case JGM_GEMM_BF16: JGM_JTR(JGM_GEMM_BF16, false); break;
      default: JGM_JTR(JGM_GEMM_BF16_REFINED, true); break;
    }
#undef JGM_JTR
    return hipGetLastError();
  });
}

// ---- FP64 J on the FP64 matrix cores -------------------------------------------------
// v_mfma_f64_16x16x4_f64: A 16 x 4 (lane l holds A[l % 16][l / 16]), B 4 x 16 (lane l holds
// B[l / 16][l % 16]), D 16 x 16 (lane l holds D[l / 16 + 4 r][l % 16] in element r).
// The k index of one instruction may map to any four cells (rows) as long as A and B
// agree: here k-step s of a 16-deep chunk takes cell (J M) or row (J^T R) cb + 4 g + s of
// lane group g = l / 16, so that a lane's four J values of a chunk are contiguous in the
// J^T R product.  A wave owns a 64 x 64 block of D as 4 x 4 tiles of 16 x 16 (FP64
// accumulators, no flush needed).
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mfma_f64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct Frag64 {
  double a[4][4], b[4][4];  // [k-step][tile]
};
// register stages of k_gemm_jm_f64: 3 (206 VGPRs with the 128 accumulator AGPRs still 1
// wave per SIMD) reads FP64 J 13 % faster than 2; 4 would spill
constexpr int kJm64Stages = 3;

__device__ __forceinline__ void mma64(const Frag64& f, f64x4 (&acc)[4][4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[t][u] = mfma_f64(f.a[s][t], f.b[s][u], acc[t][u]);
}

// Y partials Yp[split][row][P] (FP64) of J M; grid and split layout as k_gemm_jm.
__global__ __launch_bounds__(kBlock) void k_gemm_jm_f64(const double* __restrict__ J, long ld,
                                                        long R, long colpad,
                                                        const double* __restrict__ Mt, int P,
                                                        long per, int row_tiles, int nsplit,
                                                        double* __restrict__ Yp) {
  const int w = blockIdx.x;
  int split, rt;
  xcd_item(w, nsplit * row_tiles, row_tiles, split, rt);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, g = lane >> 4;
  const long row0 = (long)rt * 256 + wave * 64;
  if (row0 >= ld) return;  // whole wave outside the padded rows; no barriers follow
  const int b0 = blockIdx.y * 64;
  const long c0 = (long)split * per;
  const long cend = c0 + per < colpad ? c0 + per : colpad;
  const int nchunk = cend > c0 ? (int)((cend - c0) / kChunk) : 0;
  const double* Jw = J + row0 + li;
  const double* Mw = Mt + b0 + li;
  f64x4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[t][u] = f64x4{0.0, 0.0, 0.0, 0.0};
  // a ring of kJm64Stages chunks in registers, as in k_gemm_jm
  Frag64 f[kJm64Stages];
  auto load = [&](Frag64& h, int q) {
    const long cb = c0 + (long)(q < nchunk ? q : nchunk - 1) * kChunk;  // past the end: unused
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const long cell = cb + 4 * g + st;
#pragma unroll
      for (int t = 0; t < 4; ++t) h.a[st][t] = __builtin_nontemporal_load(Jw + cell * ld + 16 * t);
#pragma unroll
      for (int u = 0; u < 4; ++u) h.b[st][u] = Mw[cell * P + 16 * u];
    }
  };
  if (nchunk > 0) {
#pragma unroll
    for (int d = 0; d < kJm64Stages - 1; ++d) load(f[d], d);
    for (int q = 0; q < nchunk; q += kJm64Stages) {
#pragma unroll
      for (int d = 0; d < kJm64Stages; ++d) {
        load(f[(d + kJm64Stages - 1) % kJm64Stages], q + d + kJm64Stages - 1);
        __builtin_amdgcn_sched_barrier(0);
        if (q + d < nchunk) mma64(f[d], acc);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  double* out = Yp + (long)split * R * P;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long row = row0 + 16 * t + g + 4 * r;
      if (row < R) {
#pragma unroll
        for (int u = 0; u < 4; ++u) out[row * P + b0 + 16 * u + li] = acc[t][u][r];
      }
    }
}

// Y[row + b R] = sum over splits (fixed order) of Yp[split][row][b]
__global__ __launch_bounds__(kBlock) void k_gemm_jm_reduce_f64(const double* __restrict__ Yp,
                                                               int nsplit, long R, int P,
                                                               int nrhs, double* __restrict__ Y) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= R * nrhs) return;
  const long row = idx / nrhs;
  const int b = (int)(idx % nrhs);
  double a = 0.0;
  for (int s = 0; s < nsplit; ++s) a += Yp[(long)s * R * P + row * P + b];
  Y[row + (long)b * R] = a;
}

// G = J^T Rt over all rows: workgroup = 256 cells (wave: 64 cells) x 64 right-hand sides;
// A = J^T (i = cell, k = row): a lane's four rows of a chunk are one 32-byte load from its
// cell's column; B = Rt (k = row, j = right-hand side).  The next chunk's loads are in
// flight while the current one is multiplied.
__global__ __launch_bounds__(kBlock) void k_gemm_jtr_f64(const double* __restrict__ J, long ld,
                                                         long N, long nblk,
                                                         const double* __restrict__ Rt, int P,
                                                         double* __restrict__ G, long ldg,
                                                         int nrhs) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, g = lane >> 4;
  const int b0 = blockIdx.y * 64;
  const long nchunk = ld / kChunk;  // even: ld is a multiple of 64
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const long cbase = blk * kColPad + 64 * wave;
    const double* Jw = J + (cbase + li) * ld + 4 * g;
    const double* Rw = Rt + (long)(4 * g) * P + b0 + li;
    f64x4 acc[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[t][u] = f64x4{0.0, 0.0, 0.0, 0.0};
    Frag64 x, y;
    auto load = [&](Frag64& f, long rb) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const f64x4 v = __builtin_nontemporal_load((const f64x4*)(Jw + (long)16 * t * ld + rb));
#pragma unroll
        for (int st = 0; st < 4; ++st) f.a[st][t] = v[st];
      }
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int u = 0; u < 4; ++u) f.b[st][u] = Rw[(rb + st) * P + 16 * u];
    };
    load(x, 0);
    for (long q = 0; q < nchunk; q += 2) {
      load(y, (q + 1) * kChunk);
      __builtin_amdgcn_sched_barrier(0);
      mma64(x, acc);
      __builtin_amdgcn_sched_barrier(0);
      const long qn = q + 2 < nchunk ? q + 2 : nchunk - 1;  // the last reload is unused
      load(x, qn * kChunk);
      __builtin_amdgcn_sched_barrier(0);
      mma64(y, acc);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long c = cbase + 16 * t + g + 4 * r;
        if (c < N) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int b = b0 + 16 * u + li;
            if (b < nrhs) G[c + (long)b * ldg] = acc[t][u][r];
          }
        }
      }
  }
}

// J M on Xt[colpad][P] (FP64 J in HBM) -> out (R x nrhs, column-major, FP64)
hipError_t run_jm_f64(jgm_context* ctx, const double* Xt, int P, int nrhs, double* out) {
  const long R = (long)ctx->storedJ_rows, ld = (long)ctx->storedJ_ld;
  const long colpad = (long)ctx->storedJ_colpad;
  const int rhs_tiles = P / 64;
  const int row_tiles = (int)((ld + 255) / 256);
  // one pass of workgroups over the CUs (one per CU at 1 wave per SIMD): on 7168 x 65536,
  // 9 splits run 1.02 ms against 1.12 ms for 16 (448 workgroups: a second pass three
  // quarters full)
  long nsplit = std::max<long>(1, (long)ctx->cu_count / ((long)row_tiles * rhs_tiles));
  nsplit = std::min<long>(nsplit, colpad / 32);
  const long per = round_up((colpad + nsplit - 1) / nsplit, 2 * kChunk);
  hipError_t e = ctx->partial.ensure(sizeof(double) * (size_t)nsplit * R * P);
  if (e != hipSuccess) return e;
  {
    ProfScope ps(ctx, "storedj_jm_fp64");
    hipLaunchKernelGGL(k_gemm_jm_f64, dim3((unsigned)(nsplit * row_tiles), (unsigned)rhs_tiles),
                       dim3(kBlock), 0, ctx->stream, ctx->storedJ.as<const double>(), ld, R,
                       colpad, Xt, P, per, row_tiles, (int)nsplit, ctx->partial.as<double>());
  }
  hipLaunchKernelGGL(k_gemm_jm_reduce_f64, dim3(blocks_for(R * (long)nrhs)), dim3(kBlock), 0,
                     ctx->stream, ctx->partial.as<const double>(), (int)nsplit, R, P, nrhs, out);
  return hipGetLastError();
}

// J^T R on Xt[ld][P] (FP64 J in HBM) -> out (N x nrhs, column-major, FP64)
hipError_t run_jtr_f64(jgm_context* ctx, const double* Xt, int P, int nrhs, double* out) {
  const long N = (long)ctx->storedJ_cols, ld = (long)ctx->storedJ_ld;
  const long nblk = (long)ctx->storedJ_colpad / kColPad;
  ProfScope ps(ctx, "storedj_jtr_fp64");
  hipLaunchKernelGGL(k_gemm_jtr_f64, dim3((unsigned)std::min<long>(nblk, 65535), (unsigned)(P / 64)),
                     dim3(kBlock), 0, ctx->stream, ctx->storedJ.as<const double>(), ld, N, nblk,
                     Xt, P, out, N, nrhs);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_storedj_apply_f64(jgm_context* ctx, int transpose, const double* in,
                                    size_t nrhs, double* out) {
  const long R = (long)ctx->storedJ_rows, N = (long)ctx->storedJ_cols;
  const long ld = (long)ctx->storedJ_ld, colpad = (long)ctx->storedJ_colpad;
  const int P = (int)round_up((long)nrhs, 64);
  const long K = transpose ? R : N, Kpad = transpose ? ld : colpad;
  hipError_t e = ctx->coef32.ensure(sizeof(double) * (size_t)Kpad * P);
  if (e != hipSuccess) return e;
  double* Xt = ctx->coef32.as<double>();
  hipLaunchKernelGGL(k_transpose_pad<double>, dim3(blocks_for(Kpad * P)), dim3(kBlock), 0,
                     ctx->stream, in, K, Kpad, (int)nrhs, P, Xt);
  return transpose ? run_jtr_f64(ctx, Xt, P, (int)nrhs, out)
                   : run_jm_f64(ctx, Xt, P, (int)nrhs, out);
}

hipError_t launch_storedj_apply(jgm_context* ctx, int gemm_precision, int transpose,
                                const float* in, size_t nrhs, float* out) {
  const long R = (long)ctx->storedJ_rows, N = (long)ctx->storedJ_cols;
  const long ld = (long)ctx->storedJ_ld, colpad = (long)ctx->storedJ_colpad;
  const int P = (int)round_up((long)nrhs, 64);
  const long K = transpose ? R : N, Kpad = transpose ? ld : colpad;
  hipError_t e = ctx->coef32.ensure(sizeof(float) * (size_t)Kpad * P);
  if (e != hipSuccess) return e;
  float* Xt = ctx->coef32.as<float>();
  hipLaunchKernelGGL(k_transpose_pad<float>, dim3(blocks_for(Kpad * P)), dim3(kBlock), 0,
                     ctx->stream, in, K, Kpad, (int)nrhs, P, Xt);
  return transpose ? run_jtr(ctx, gemm_precision, Xt, P, (int)nrhs, out)
                   : run_jm(ctx, gemm_precision, Xt, P, (int)nrhs, out);
}

hipError_t launch_storedj_forward(jgm_context* ctx, int gemm_precision, const double* model,
                                  double* out, double* raw_out) {
  return launch_storedj_forward_batch(ctx, gemm_precision, 1, model, 0, out, 0, raw_out, 0);
}

namespace {

// J M of P right-hand sides, FP32 modes on an FP32 J or FP64 on an FP64 J
hipError_t run_jm_t(jgm_context* ctx, int gemm_precision, const float* Xt, int P, int nrhs,
                    float* out) {
  return run_jm(ctx, gemm_precision, Xt, P, nrhs, out);
}
hipError_t run_jm_t(jgm_context* ctx, int, const double* Xt, int P, int nrhs, double* out) {
  return run_jm_f64(ctx, Xt, P, nrhs, out);
}
hipError_t run_jtr_t(jgm_context* ctx, int gemm_precision, const float* Xt, int P, int nrhs,
                     float* out) {
  return run_jtr(ctx, gemm_precision, Xt, P, nrhs, out);
}
hipError_t run_jtr_t(jgm_context* ctx, int, const double* Xt, int P, int nrhs, double* out) {
  return run_jtr_f64(ctx, Xt, P, nrhs, out);
}

// Up to 64 models per J M product (the product's right-hand-side tile is 64 wide, so
// a batch costs what one model costs); larger batches take several products.  T is the
// element type of the stored J and of the products' operands.
template <class T>
hipError_t forward_batch_impl(jgm_context* ctx, int gemm_precision, size_t nmodels,
                              const double* models, size_t model_ld, double* out, size_t out_ld,
                              double* raw_out, size_t raw_ld) {
  const long R = (long)ctx->storedJ_rows, N = (long)ctx->storedJ_cols;
  const long colpad = (long)ctx->storedJ_colpad;
  const int kind = ctx->storedJ_kind;
  const int S = (int)ctx->S;
  const int L = has_background(kind) ? (int)ctx->bgthick.size() : 0;
  const int P = 64;
  hipError_t e = ctx->coef32.ensure(sizeof(T) * (size_t)colpad * P);
  if (e != hipSuccess) return e;
  const size_t nb_max = std::min<size_t>(nmodels, (size_t)P);
  if ((e = ctx->storedJ_vec.ensure(sizeof(T) * (size_t)std::max(R * (long)nb_max, N))) !=
      hipSuccess)
    return e;
  T* Xt = ctx->coef32.as<T>();
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  for (size_t b0 = 0; b0 < nmodels; b0 += P) {
    const int nb = (int)std::min<size_t>(nmodels - b0, (size_t)P);
    const double* m = models + b0 * model_ld;
    hipLaunchKernelGGL(k_vec_to_rhs<T>, dim3(blocks_for(colpad * P)), dim3(kBlock), 0,
                       ctx->stream, m, (long)model_ld, nb, N, colpad, P, Xt);
    T* Y = ctx->storedJ_vec.as<T>();
    if ((e = run_jm_t(ctx, gemm_precision, Xt, P, nb, Y)) != hipSuccess) return e;
    for (int b = 0; b < nb; ++b)
      hipLaunchKernelGGL(k_storedj_finish<T>, dim3(blocks_for(S)), dim3(kBlock), 0, ctx->stream,
                         Y + (size_t)b * R, S, kind, ctx->bgT.as<const double>(),
                         m + b * model_ld + (L > 0 ? N : 0), L, field_dir(ctx), ctx->transform,
                         tf_bx, tf_by, tf_bz, out + (b0 + b) * out_ld,
                         raw_out ? raw_out + (b0 + b) * raw_ld : nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

template <class T>
hipError_t gradient_batch_impl(jgm_context* ctx, int gemm_precision, size_t nsets,
                               const double* raw_misfits, size_t misfit_ld, double* grads,
                               size_t grad_ld) {
  const long R = (long)ctx->storedJ_rows, N = (long)ctx->storedJ_cols;
  const long ld = (long)ctx->storedJ_ld;
  const int kind = ctx->storedJ_kind;
  const int P = 64;
  hipError_t e = ctx->coef32.ensure(sizeof(T) * (size_t)ld * P);
  if (e != hipSuccess) return e;
  const size_t nb_max = std::min<size_t>(nsets, (size_t)P);
  if ((e = ctx->storedJ_vec.ensure(sizeof(T) * (size_t)std::max(R, N * (long)nb_max))) !=
      hipSuccess)
    return e;
  T* Xt = ctx->coef32.as<T>();
  for (size_t b0 = 0; b0 < nsets; b0 += P) {
    const int nb = (int)std::min<size_t>(nsets - b0, (size_t)P);
    const double* rm = raw_misfits + b0 * misfit_ld;
    hipLaunchKernelGGL(k_rm_to_compact<T>, dim3(blocks_for(ld * P)), dim3(kBlock), 0,
                       ctx->stream, rm, (long)misfit_ld, nb, (int)ctx->S, kind, ld, P, Xt);
    T* G = ctx->storedJ_vec.as<T>();
    if ((e = run_jtr_t(ctx, gemm_precision, Xt, P, nb, G)) != hipSuccess) return e;
    for (int b = 0; b < nb; ++b) {
      double* g = grads + (b0 + b) * grad_ld;
      hipLaunchKernelGGL(k_scale_to_f64<T>, dim3(blocks_for(N)), dim3(kBlock), 0, ctx->stream,
                         G + (size_t)b * N, N, 2.0, g);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if (has_background(kind))
        e = launch_background_gradient(ctx, kind, rm + b * misfit_ld, g + N);
      else if (kind == JGM_KIND_MAGNETIC_GRADIENT)  // zero background columns
        e = launch_fill_zero(ctx, g + N, ctx->bgthick.size());
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

}  // namespace

hipError_t launch_storedj_forward_batch(jgm_context* ctx, int gemm_precision, size_t nmodels,
                                        const double* models, size_t model_ld, double* out,
                                        size_t out_ld, double* raw_out, size_t raw_ld) {
  if (ctx->storedJ_fp64)
    return forward_batch_impl<double>(ctx, gemm_precision, nmodels, models, model_ld, out,
                                      out_ld, raw_out, raw_ld);
  return forward_batch_impl<float>(ctx, gemm_precision, nmodels, models, model_ld, out, out_ld,
                                   raw_out, raw_ld);
}

hipError_t launch_storedj_gradient(jgm_context* ctx, int gemm_precision,
                                   const double* raw_misfit, double* grad) {
  return launch_storedj_gradient_batch(ctx, gemm_precision, 1, raw_misfit, 0, grad, 0);
}

hipError_t launch_storedj_gradient_batch(jgm_context* ctx, int gemm_precision, size_t nsets,
                                         const double* raw_misfits, size_t misfit_ld,
                                         double* grads, size_t grad_ld) {
  if (ctx->storedJ_fp64)
    return gradient_batch_impl<double>(ctx, gemm_precision, nsets, raw_misfits, misfit_ld,
                                       grads, grad_ld);
  return gradient_batch_impl<float>(ctx, gemm_precision, nsets, raw_misfits, misfit_ld, grads,
                                    grad_ld);
}

}  // namespace jgm
