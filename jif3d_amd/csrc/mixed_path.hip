// Mixed-precision path.
//
// Split per site: cells inside the site's near box (every cell closer than q0
// cell half-diagonals lies inside it) use the FP64 closed form; all other cells use
// the FP32 multipole expansion of the uniform prism about its centre, monopole plus
// quadrupole:
//
//   Phi/V = 1/r + Q/r^5,  Q = Bx X^2 + By Y^2 + Bz Z^2,  B_m = (h_m^2 - hbar^2)/2,
//   hbar^2 = (hx^2 + hy^2 + hz^2)/3, (X, Y, Z) = site - cell centre, h = half sizes,
//
// gz = G rho V dPhi/dZ and U_ij = G rho V d2Phi/dx_i dx_j (the octupole vanishes by
// symmetry; the neglected hexadecapole is O((h/d)^4) relative).  Cubes have B = 0.
// Accumulation per site (forward) and per cell (adjoint) runs in FP32 within a
// 4x4x4-cell tile and in FP64 across tiles.
//
// Layouts in HBM
//   cell table   FP32, QUAD: [c][8] = {w, Bx, By, Bz, Bx+By, Bx+Bz, By+Bz, 0}; else [c] = {w}
//                w = G rho V (forward) or G V (adjoint)
//   far partials forward [chunk][NC][S] FP64; adjoint [chunk][N] FP64
//   near boxes   int32 [S][6]; per-column CSR of sites for the adjoint near pass
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
constexpr int kTile = 4;  // far-field tile edge (cells)
constexpr int kTargetWaves = 16384;

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__global__ void k_fill_zero(double* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Far field: FP32 multipole of a prism of unit G*rho*V.  Outputs in the kernel
// component order of MASK (gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz).
// ---------------------------------------------------------------------------
template <int MASK, bool QUAD>
__device__ __forceinline__ void multipole_f32(float X, float Y, float Z, float X2, float Y2,
                                              float XY, const float* B, float* T) {
  const float Z2 = Z * Z;
  const float r2 = X2 + Y2 + Z2;
  const float s = __builtin_amdgcn_rsqf(r2);
  const float s2 = s * s;
  const float s3 = s2 * s;
  const float s5 = s3 * s2;
  int k = 0;
  if constexpr (!QUAD) {
    const float t5 = 3.0f * s5;
    if constexpr ((MASK & kGz) != 0) T[k++] = -Z * s3;
    if constexpr ((MASK & kDiag) != 0) {
      T[k++] = fmaf(X2, t5, -s3);
      T[k++] = fmaf(Y2, t5, -s3);
      T[k++] = fmaf(Z2, t5, -s3);
    }
    if constexpr ((MASK & kOff) != 0) {
      T[k++] = XY * t5;
      T[k++] = (X * Z) * t5;
      T[k++] = (Y * Z) * t5;
    }
  } else {
    const float Bx = B[0], By = B[1], Bz = B[2];
    const float s7 = s5 * s2;
    const float s9 = s7 * s2;
    const float Q = fmaf(Bx, X2, fmaf(By, Y2, Bz * Z2));
    const float base = fmaf(-5.0f * Q, s7, -s3);        // -s3 - 5 Q s7
    const float P = fmaf(35.0f * Q, s9, 3.0f * s5);      // 3 s5 + 35 Q s9
    const float t2s5 = 2.0f * s5;
    if constexpr ((MASK & kGz) != 0) T[k++] = Z * fmaf(Bz, t2s5, base);
    if constexpr ((MASK & kDiag) != 0) {
      const float m20s7 = -20.0f * s7;
      T[k++] = fmaf(X2, fmaf(Bx, m20s7, P), fmaf(Bx, t2s5, base));
      T[k++] = fmaf(Y2, fmaf(By, m20s7, P), fmaf(By, t2s5, base));
      T[k++] = fmaf(Z2, fmaf(Bz, m20s7, P), fmaf(Bz, t2s5, base));
    }
    if constexpr ((MASK & kOff) != 0) {
      const float m10s7 = -10.0f * s7;
      T[k++] = XY * fmaf(B[3], m10s7, P);
      T[k++] = (X * Z) * fmaf(B[4], m10s7, P);
      T[k++] = (Y * Z) * fmaf(B[5], m10s7, P);
    }
  }
}

// Cell table: w = G * rho * V (rho = 1 when rho is null) and the quadrupole constants.
template <bool QUAD>
__global__ __launch_bounds__(kBlock) void k_cell_table(const double* __restrict__ rho,
                                                       const double* __restrict__ xsz,
                                                       const double* __restrict__ ysz,
                                                       const double* __restrict__ zsz, int nx,
                                                       int ny, int nz, float* __restrict__ table) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)nx * ny * nz;
  if (c >= n) return;
  const int k = (int)(c % nz);
  const long t = c / nz;
  const int j = (int)(t % ny);
  const int i = (int)(t / ny);
  const double dx = xsz[i], dy = ysz[j], dz = zsz[k];
  const double V = dx * dy * dz;
  const double w = kGravConst * (rho ? rho[c] : 1.0) * V;
  if constexpr (!QUAD) {
    table[c] = (float)w;
  } else {
    const double hx2 = 0.25 * dx * dx, hy2 = 0.25 * dy * dy, hz2 = 0.25 * dz * dz;
    const double hb = (hx2 + hy2 + hz2) / 3.0;
    const double Bx = 0.5 * (hx2 - hb), By = 0.5 * (hy2 - hb), Bz = 0.5 * (hz2 - hb);
    float* o = table + c * 8;
    o[0] = (float)w;
    o[1] = (float)Bx;
    o[2] = (float)By;
    o[3] = (float)Bz;
    o[4] = (float)(Bx + By);
    o[5] = (float)(Bx + Bz);
    o[6] = (float)(By + Bz);
    o[7] = 0.0f;
  }
}

struct FarGeom {
  const double* xcen;
  const double* ycen;
  const double* zcen;
  const float* offx;
  const float* offy;
  const float* offz;
  int nx, ny, nz;
  int ntx, nty, ntz;
};

// Forward far field: one lane per site; each wave walks a chunk of 4x4x4 tiles.
// Cells inside the lane's near box contribute nothing here (FP64 near pass).
template <int MASK, bool QUAD>
__global__ __launch_bounds__(kBlock) void k_far_forward(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const int* __restrict__ box, FarGeom g, const float* __restrict__ table, long ntiles,
    long tiles_per_chunk, double* __restrict__ farpart) {
  constexpr int NC = Ncomp<MASK>::value;
  constexpr int TW = QUAD ? 8 : 1;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = s < S ? s : S - 1;
  const double px = sx[sl], py = sy[sl], pz = sz[sl];
  const int b_i0 = box[sl * 6 + 0], b_i1 = box[sl * 6 + 1];
  const int b_j0 = box[sl * 6 + 2], b_j1 = box[sl * 6 + 3];
  const int b_k0 = box[sl * 6 + 4], b_k1 = box[sl * 6 + 5];
  const long t0 = (long)blockIdx.y * tiles_per_chunk;
  const long t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;
  double acc64[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc64[m] = 0.0;
  const int ny = g.ny, nz = g.nz;
  for (long t = t0; t < t1; ++t) {
    const int tk = (int)(t % g.ntz);
    const long tt = t / g.ntz;
    const int tj = (int)(tt % g.nty);
    const int ti = (int)(tt / g.nty);
    const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * kTile;
    const int ni = min(kTile, g.nx - i0), nj = min(kTile, ny - j0), nk = min(kTile, nz - k0);
    // site relative to the tile's first cell centre, rounded once to FP32
    const float pxr = (float)(px - g.xcen[i0]);
    const float pyr = (float)(py - g.ycen[j0]);
    const float pzr = (float)(pz - g.zcen[k0]);
    const bool hit = !(b_i1 < i0 || b_i0 > i0 + ni - 1 || b_j1 < j0 || b_j0 > j0 + nj - 1 ||
                       b_k1 < k0 || b_k0 > k0 + nk - 1);
    float acc[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[m] = 0.0f;
    if (!__any(hit)) {
      for (int ii = 0; ii < ni; ++ii) {
        const float X = pxr - g.offx[i0 + ii];
        const float X2 = X * X;
        for (int jj = 0; jj < nj; ++jj) {
          const float Y = pyr - g.offy[j0 + jj];
          const float Y2 = Y * Y, XY = X * Y;
          const float* tb = table + (((long)(i0 + ii) * ny + (j0 + jj)) * nz + k0) * TW;
          for (int kk = 0; kk < nk; ++kk) {
            const float Z = pzr - g.offz[k0 + kk];
            const float* cw = tb + kk * TW;
            float T[NC];
            multipole_f32<MASK, QUAD>(X, Y, Z, X2, Y2, XY, cw + 1, T);
            const float w = cw[0];
#pragma unroll
            for (int m = 0; m < NC; ++m) acc[m] = fmaf(w, T[m], acc[m]);
          }
        }
      }
    } else {
      for (int ii = 0; ii < ni; ++ii) {
        const float X = pxr - g.offx[i0 + ii];
        const float X2 = X * X;
        const bool inx = (i0 + ii >= b_i0) && (i0 + ii <= b_i1);
        for (int jj = 0; jj < nj; ++jj) {
          const float Y = pyr - g.offy[j0 + jj];
          const float Y2 = Y * Y, XY = X * Y;
          const bool inxy = inx && (j0 + jj >= b_j0) && (j0 + jj <= b_j1);
          const float* tb = table + (((long)(i0 + ii) * ny + (j0 + jj)) * nz + k0) * TW;
          for (int kk = 0; kk < nk; ++kk) {
            const bool near = inxy && (k0 + kk >= b_k0) && (k0 + kk <= b_k1);
            // a near cell is evaluated at a harmless far point and weighted by zero
            const float Z = near ? 1.0e6f : pzr - g.offz[k0 + kk];
            const float* cw = tb + kk * TW;
            float T[NC];
            multipole_f32<MASK, QUAD>(X, Y, Z, X2, Y2, XY, cw + 1, T);
            const float w = near ? 0.0f : cw[0];
#pragma unroll
            for (int m = 0; m < NC; ++m) acc[m] = fmaf(w, T[m], acc[m]);
          }
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NC; ++m) acc64[m] += (double)acc[m];
  }
  if (s < S) {
#pragma unroll
    for (int m = 0; m < NC; ++m) farpart[((long)blockIdx.y * NC + m) * S + s] = acc64[m];
  }
}

// Forward near field: one block per site, FP64 closed form over the corners of
// the site's near box with weights w = sum sigma rho over the box's cells.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_near_forward(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const int* __restrict__ box, const double* __restrict__ Xc, const double* __restrict__ Yc,
    const double* __restrict__ Zc, const double* __restrict__ rho, int nx, int ny, int nz,
    double* __restrict__ nearout) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x;
  const int i0 = box[s * 6 + 0], i1 = box[s * 6 + 1];
  const int j0 = box[s * 6 + 2], j1 = box[s * 6 + 3];
  const int k0 = box[s * 6 + 4], k1 = box[s * 6 + 5];
  double acc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = 0.0;
  if (i0 <= i1 && j0 <= j1 && k0 <= k1) {
    const double px = sx[s], py = sy[s], pz = sz[s];
    const int ci = i1 - i0 + 2, cj = j1 - j0 + 2, ck = k1 - k0 + 2;
    const int nq = ci * cj * ck;
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
      const int K = k0 + q % ck;
      const int J = j0 + (q / ck) % cj;
      const int I = i0 + q / (ck * cj);
      double w = 0.0;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            const int i = I - a, j = J - b, k = K - c;
            if (i < i0 || i > i1 || j < j0 || j > j1 || k < k0 || k > k1) continue;
            const double v = rho[((long)i * ny + j) * nz + k];
            w += ((a + b + c) & 1) ? v : -v;
          }
      double t[NC];
      corner_terms_f64<MASK>(px - Xc[I], py - Yc[J], pz - Zc[K], t);
#pragma unroll
      for (int m = 0; m < NC; ++m) acc[m] = fma(w, t[m], acc[m]);
    }
  }
  __shared__ double red[kBlock];
  double sc[NC];
  comp_scales<MASK>(sc);
  for (int m = 0; m < NC; ++m) {
    red[threadIdx.x] = acc[m];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) nearout[(long)s * NC + m] = red[0] * sc[m];
    __syncthreads();
  }
}

template <int MASK>
__global__ __launch_bounds__(kBlock) void k_mixed_forward_finish(
    const double* __restrict__ farpart, int nchunks, const double* __restrict__ nearout, int S,
    const double* __restrict__ bgT, const double* __restrict__ bgvals, int L, int kind,
    FieldDir fd, int tf, double tfbx, double tfby, double tfbz, double* __restrict__ out,
    double* __restrict__ raw_out) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double acc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = nearout[(long)s * NC + m];
  for (int c = 0; c < nchunks; ++c)
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[m] += farpart[((long)c * NC + m) * S + s];
  double u[7];
  spread_components<MASK>(acc, u);
  finish_site(u, s, bgT, bgvals, L, kind, fd, tf, tfbx, tfby, tfbz, out, raw_out);
}

// Physical coefficients of the adjoint (before sign*G), FP64 and FP32 copies.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_phys_coef(const double* __restrict__ rm, int S,
                                                      int kind, FieldDir fd,
                                                      double* __restrict__ c64,
                                                      float* __restrict__ c32) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double c7[7], c[NC];
  site_coefficients(rm, s, kind, fd, c7);
  select_components<MASK>(c7, c);
  for (int m = 0; m < NC; ++m) {
    c64[(long)s * NC + m] = c[m];
    c32[(long)s * NC + m] = (float)c[m];
  }
}

// Adjoint far field: one lane per cell of a 4x4x4 tile, the wave walks a chunk
// of sites (wave-uniform); sites whose near box holds the cell are skipped.
template <int MASK, bool QUAD>
__global__ __launch_bounds__(kBlock) void k_far_adjoint(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const float* __restrict__ coef, const int* __restrict__ box, int S, int sites_per_chunk,
    FarGeom g, const float* __restrict__ table, long ntiles, double* __restrict__ gpart) {
  constexpr int NC = Ncomp<MASK>::value;
  constexpr int TW = QUAD ? 8 : 1;
  const long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (t >= ntiles) return;
  const int tk = (int)(t % g.ntz);
  const long tt = t / g.ntz;
  const int tj = (int)(tt % g.nty);
  const int ti = (int)(tt / g.nty);
  const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * kTile;
  const int ni = min(kTile, g.nx - i0), nj = min(kTile, g.ny - j0), nk = min(kTile, g.nz - k0);
  const int ii = lane >> 4, jj = (lane >> 2) & 3, kk = lane & 3;
  const bool valid = ii < ni && jj < nj && kk < nk;
  const int i = i0 + (valid ? ii : 0), j = j0 + (valid ? jj : 0), k = k0 + (valid ? kk : 0);
  const long c = ((long)i * g.ny + j) * g.nz + k;
  const float dx = g.offx[i], dy = g.offy[j], dz = g.offz[k];
  float B[6] = {0, 0, 0, 0, 0, 0};
  if constexpr (QUAD) {
    for (int q = 0; q < 6; ++q) B[q] = table[c * 8 + 1 + q];
  }
  const double xr = g.xcen[i0], yr = g.ycen[j0], zr = g.zcen[k0];
  const int s0 = blockIdx.y * sites_per_chunk;
  const int s1 = s0 + sites_per_chunk < S ? s0 + sites_per_chunk : S;
  double acc64 = 0.0;
  float acc = 0.0f;
  int nacc = 0;
  for (int s = s0; s < s1; ++s) {
    const float pxr = (float)(sx[s] - xr);
    const float pyr = (float)(sy[s] - yr);
    const float pzr = (float)(sz[s] - zr);
    const int* bx = box + s * 6;
    const bool tile_hit = !(bx[1] < i0 || bx[0] > i0 + ni - 1 || bx[3] < j0 ||
                            bx[2] > j0 + nj - 1 || bx[5] < k0 || bx[4] > k0 + nk - 1);
    bool near = false;
    if (tile_hit)
      near = i >= bx[0] && i <= bx[1] && j >= bx[2] && j <= bx[3] && k >= bx[4] && k <= bx[5];
    const float X = pxr - dx, Y = pyr - dy;
    const float Z = near ? 1.0e6f : pzr - dz;
    float T[NC];
    multipole_f32<MASK, QUAD>(X, Y, Z, X * X, Y * Y, X * Y, B, T);
    const float* cs = coef + (long)s * NC;
    float v = 0.0f;
#pragma unroll
    for (int m = 0; m < NC; ++m) v = fmaf(cs[m], T[m], v);
    acc += near ? 0.0f : v;
    if (++nacc == 64) {
      acc64 += (double)acc;
      acc = 0.0f;
      nacc = 0;
    }
  }
  acc64 += (double)acc;
  if (valid) {
    const double w = (double)table[c * TW];  // G * V
    gpart[(long)blockIdx.y * ((long)g.nx * g.ny * g.nz) + c] = 2.0 * w * acc64;
  }
}

// Adjoint near field: one block per listed column (i, j); threads over k; each
// cell gathers its near sites (those whose box holds it) in ascending site order
// with the FP64 8-corner closed form.
template <int MASK>
__global__ __launch_bounds__(64) void k_near_adjoint(
    const int* __restrict__ col_ptr, const int* __restrict__ col_sites,
    const int* __restrict__ col_ids, const int* __restrict__ col_kmax,
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const int* __restrict__ box, const double* __restrict__ coef, const double* __restrict__ xlo,
    const double* __restrict__ ylo, const double* __restrict__ zlo,
    const double* __restrict__ xsz, const double* __restrict__ ysz,
    const double* __restrict__ zsz, int ny, int nz, double* __restrict__ gnear) {
  constexpr int NC = Ncomp<MASK>::value;
  const int col = col_ids[blockIdx.x];
  const int i = col / ny, j = col % ny;
  const int p0 = col_ptr[blockIdx.x], p1 = col_ptr[blockIdx.x + 1];
  const int kmax = col_kmax[blockIdx.x];
  for (int k = threadIdx.x; k <= kmax; k += blockDim.x) {
    double acc = 0.0;
    for (int p = p0; p < p1; ++p) {
      const int s = col_sites[p];
      if (k < box[s * 6 + 4] || k > box[s * 6 + 5]) continue;
      double v[NC];
      box_terms_f64<MASK>(sx[s], sy[s], sz[s], xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
      const double* cs = coef + (long)s * NC;
#pragma unroll
      for (int m = 0; m < NC; ++m) acc = fma(cs[m], v[m], acc);
    }
    gnear[((long)i * ny + j) * nz + k] = 2.0 * acc;
  }
}

__global__ __launch_bounds__(kBlock) void k_adjoint_mixed_finish(const double* __restrict__ gpart,
                                                                 int nchunks,
                                                                 const double* __restrict__ gnear,
                                                                 long n,
                                                                 double* __restrict__ grad) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double a = gnear[c];
  for (int q = 0; q < nchunks; ++q) a += gpart[(long)q * n + c];
  grad[c] = a;
}

int mask_of(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return kMaskScalar;
    case JGM_KIND_SCALAR_TENSOR: return kMaskAll;
    default: return kMaskTensor;
  }
}

FarGeom far_geom(const jgm_context* ctx) {
  FarGeom g;
  g.xcen = ctx->near.xcen.as<const double>();
  g.ycen = ctx->near.ycen.as<const double>();
  g.zcen = ctx->near.zcen.as<const double>();
  g.offx = ctx->near.offx.as<const float>();
  g.offy = ctx->near.offy.as<const float>();
  g.offz = ctx->near.offz.as<const float>();
  g.nx = (int)ctx->nx;
  g.ny = (int)ctx->ny;
  g.nz = (int)ctx->nz;
  g.ntx = (g.nx + kTile - 1) / kTile;
  g.nty = (g.ny + kTile - 1) / kTile;
  g.ntz = (g.nz + kTile - 1) / kTile;
  return g;
}

template <class T>
hipError_t upload_vec(DeviceBuffer& b, const std::vector<T>& v, hipStream_t st) {
  hipError_t e = b.ensure(sizeof(T) * std::max<size_t>(1, v.size()));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpyAsync(b.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st);
}

// Index range of cells whose centre lies strictly within R of p along one axis.
void axis_range(const std::vector<double>& cen, double p, double R, int* lo, int* hi) {
  auto a = std::upper_bound(cen.begin(), cen.end(), p - R);
  auto b = std::lower_bound(cen.begin(), cen.end(), p + R);
  *lo = (int)(a - cen.begin());
  *hi = (int)(b - cen.begin()) - 1;
}

template <int MASK, bool QUAD>
hipError_t forward_mixed_impl(jgm_context* ctx, int kind, const double* model, double* out,
                              double* raw_out) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const FarGeom g = far_geom(ctx);
  const long ntiles = (long)g.ntx * g.nty * g.ntz;
  hipError_t e;
  int nchunks = 0;
  if (N > 0) {
    const int site_blocks = (S + kBlock - 1) / kBlock;
    const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
    long want = (kTargetWaves + waves_per_chunk - 1) / waves_per_chunk;
    want = std::max<long>(1, std::min<long>(want, ntiles));
    const long per = (ntiles + want - 1) / want;
    nchunks = (int)((ntiles + per - 1) / per);
    if ((e = ctx->weights.ensure(sizeof(float) * (size_t)N * (QUAD ? 8 : 1))) != hipSuccess)
      return e;
    if ((e = ctx->farpart.ensure(sizeof(double) * (size_t)nchunks * NC * S)) != hipSuccess)
      return e;
    if ((e = ctx->nearpart.ensure(sizeof(double) * (size_t)S * NC)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_cell_table<QUAD>), dim3(blocks_for(N)), dim3(kBlock), 0, st, model,
                       ctx->d_xsz.as<const double>(), ctx->d_ysz.as<const double>(),
                       ctx->d_zsz.as<const double>(), g.nx, g.ny, g.nz, ctx->weights.as<float>());
    {
      ProfScope ps(ctx, MASK == kMaskScalar ? "far_forward_fp32_gz"
                        : MASK == kMaskAll  ? "far_forward_fp32_gz_ftg"
                                            : "far_forward_fp32_ftg");
      hipLaunchKernelGGL((k_far_forward<MASK, QUAD>), dim3(site_blocks, nchunks), dim3(kBlock), 0,
                         st, ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),
                         ctx->d_sz.as<double>(), S, ctx->near.box.as<const int>(), g,
                         ctx->weights.as<const float>(), ntiles, per, ctx->farpart.as<double>());
    }
    {
      ProfScope ps(ctx, "near_forward_fp64");
      hipLaunchKernelGGL((k_near_forward<MASK>), dim3(S), dim3(kBlock), 0, st,
                         ctx->d_sx.as<double>(), ctx->d_sy.as<double>(), ctx->d_sz.as<double>(),
                         ctx->near.box.as<const int>(), ctx->d_Xc.as<double>(),
                         ctx->d_Yc.as<double>(), ctx->d_Zc.as<double>(), model, g.nx, g.ny, g.nz,
                         ctx->nearpart.as<double>());
    }
  } else {
    if ((e = ctx->nearpart.ensure(sizeof(double) * (size_t)S * NC)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for((long)S * NC)), dim3(kBlock), 0, st,
                       ctx->nearpart.as<double>(), (size_t)S * NC);
  }
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  hipLaunchKernelGGL((k_mixed_forward_finish<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st,
                     ctx->farpart.as<const double>(), nchunks, ctx->nearpart.as<const double>(), S,
                     ctx->bgT.as<const double>(), model + N, L, kind, field_dir(ctx),
                     ctx->transform, tf_bx, tf_by, tf_bz, out, raw_out);
  return hipGetLastError();
}

template <int MASK, bool QUAD>
hipError_t adjoint_mixed_impl(jgm_context* ctx, int kind, const double* rm, double* grad) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const FarGeom g = far_geom(ctx);
  const long ntiles = (long)g.ntx * g.nty * g.ntz;
  hipError_t e;
  if (N > 0) {
    if ((e = ctx->coef.ensure(sizeof(double) * (size_t)S * NC)) != hipSuccess) return e;
    if ((e = ctx->coef32.ensure(sizeof(float) * (size_t)S * NC)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_phys_coef<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st, rm, S, kind,
                       field_dir(ctx), ctx->coef.as<double>(), ctx->coef32.as<float>());
    if (!ctx->near.adj_table_valid) {
      if ((e = ctx->near.adj_table.ensure(sizeof(float) * (size_t)N * 8)) != hipSuccess) return e;
      hipLaunchKernelGGL((k_cell_table<QUAD>), dim3(blocks_for(N)), dim3(kBlock), 0, st,
                         (const double*)nullptr, ctx->d_xsz.as<const double>(),
                         ctx->d_ysz.as<const double>(), ctx->d_zsz.as<const double>(), g.nx, g.ny,
                         g.nz, ctx->near.adj_table.as<float>());
      ctx->near.adj_table_valid = true;
    }
    const long tile_waves = ntiles;
    int nchunks = (int)std::min<long>(
        S, std::max<long>(1, (kTargetWaves + tile_waves - 1) / tile_waves));
    const int per = (S + nchunks - 1) / nchunks;
    nchunks = (S + per - 1) / per;
    if ((e = ctx->hpart.ensure(sizeof(double) * (size_t)nchunks * N)) != hipSuccess) return e;
    if ((e = ctx->hsum.ensure(sizeof(double) * (size_t)N)) != hipSuccess) return e;
    {
      ProfScope ps(ctx, MASK == kMaskScalar ? "far_adjoint_fp32_gz"
                        : MASK == kMaskAll  ? "far_adjoint_fp32_gz_ftg"
                                            : "far_adjoint_fp32_ftg");
      hipLaunchKernelGGL((k_far_adjoint<MASK, QUAD>), dim3(blocks_for(ntiles * 64), nchunks),
                         dim3(kBlock), 0, st, ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),
                         ctx->d_sz.as<double>(), ctx->coef32.as<const float>(),
                         ctx->near.box.as<const int>(), S, per, g,
                         ctx->near.adj_table.as<const float>(), ntiles, ctx->hpart.as<double>());
    }
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(N)), dim3(kBlock), 0, st,
                       ctx->hsum.as<double>(), (size_t)N);
    if (ctx->near.ncols > 0) {
      ProfScope ps(ctx, "near_adjoint_fp64");
      hipLaunchKernelGGL((k_near_adjoint<MASK>), dim3((unsigned)ctx->near.ncols), dim3(64), 0, st,
                         ctx->near.col_ptr.as<const int>(), ctx->near.col_sites.as<const int>(),
                         ctx->near.col_ids.as<const int>(), ctx->near.col_kmax.as<const int>(),
                         ctx->d_sx.as<double>(), ctx->d_sy.as<double>(), ctx->d_sz.as<double>(),
                         ctx->near.box.as<const int>(), ctx->coef.as<const double>(),
                         ctx->d_xlo.as<double>(), ctx->d_ylo.as<double>(),
                         ctx->d_zlo.as<double>(), ctx->d_xsz.as<double>(),
                         ctx->d_ysz.as<double>(), ctx->d_zsz.as<double>(), g.ny, g.nz,
                         ctx->hsum.as<double>());
    }
    hipLaunchKernelGGL(k_adjoint_mixed_finish, dim3(blocks_for(N)), dim3(kBlock), 0, st,
                       ctx->hpart.as<const double>(), nchunks, ctx->hsum.as<const double>(), N,
                       grad);
  }
  if (!ctx->bgthick.empty()) {
    e = launch_background_gradient(ctx, kind, rm, grad + N);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_fill_zero(jgm_context* ctx, double* p, size_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     ctx->stream, p, n);
  return hipGetLastError();
}

// Host-side planning of the near/far split (geometry only): per-site index boxes,
// the per-column site lists of the adjoint near pass, and the far-field geometry.
hipError_t build_near_boxes(jgm_context* ctx) {
  NearBoxes& nb = ctx->near;
  const int nx = (int)ctx->nx, ny = (int)ctx->ny, nz = (int)ctx->nz;
  const size_t S = ctx->S;
  std::vector<double> xcen(nx), ycen(ny), zcen(nz);
  for (int i = 0; i < nx; ++i) xcen[i] = ctx->xc[i] + 0.5 * ctx->xs[i];
  for (int j = 0; j < ny; ++j) ycen[j] = ctx->yc[j] + 0.5 * ctx->ys[j];
  for (int k = 0; k < nz; ++k) zcen[k] = ctx->zc[k] + 0.5 * ctx->zs[k];
  double mx = 0, my = 0, mz = 0;
  for (double v : ctx->xs) mx = std::max(mx, std::fabs(v));
  for (double v : ctx->ys) my = std::max(my, std::fabs(v));
  for (double v : ctx->zs) mz = std::max(mz, std::fabs(v));
  // any cell outside the box is >= R away along some axis, and R >= q0 * its half-diagonal
  const double R = ctx->mixed_q0 * 0.5 * std::sqrt(mx * mx + my * my + mz * mz);
  nb.radius = R;
  bool cubes = true;
  const double c0 = nx ? ctx->xs[0] : 0.0;
  for (double v : ctx->xs) cubes = cubes && v == c0;
  for (double v : ctx->ys) cubes = cubes && v == c0;
  for (double v : ctx->zs) cubes = cubes && v == c0;
  nb.quad = !cubes;
  // monotone centres are required by the binary searches
  bool monotone = true;
  for (int i = 1; i < nx; ++i) monotone = monotone && xcen[i] > xcen[i - 1];
  for (int j = 1; j < ny; ++j) monotone = monotone && ycen[j] > ycen[j - 1];
  for (int k = 1; k < nz; ++k) monotone = monotone && zcen[k] > zcen[k - 1];
  if (!monotone) return hipErrorInvalidValue;
  std::vector<int> box(std::max<size_t>(1, S) * 6);
  std::vector<int> count((size_t)nx * ny + 1, 0);
  size_t total_cells = 0;
  for (size_t s = 0; s < S; ++s) {
    int lo[3], hi[3];
    axis_range(xcen, ctx->hsx[s], R, &lo[0], &hi[0]);
    axis_range(ycen, ctx->hsy[s], R, &lo[1], &hi[1]);
    axis_range(zcen, ctx->hsz[s], R, &lo[2], &hi[2]);
    bool empty = false;
    for (int a = 0; a < 3; ++a) empty = empty || lo[a] > hi[a];
    if (empty) {
      lo[0] = lo[1] = lo[2] = 1;
      hi[0] = hi[1] = hi[2] = 0;
    } else {
      total_cells += (size_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
      for (int i = lo[0]; i <= hi[0]; ++i)
        for (int j = lo[1]; j <= hi[1]; ++j) count[(size_t)i * ny + j]++;
    }
    for (int a = 0; a < 3; ++a) {
      box[s * 6 + 2 * a] = lo[a];
      box[s * 6 + 2 * a + 1] = hi[a];
    }
  }
  nb.total_near_cells = total_cells;
  std::vector<int> col_ids, col_ptr(1, 0), col_kmax;
  std::vector<int> slot((size_t)nx * ny, -1);
  for (size_t col = 0; col < (size_t)nx * ny; ++col) {
    if (count[col] == 0) continue;
    slot[col] = (int)col_ids.size();
    col_ids.push_back((int)col);
    col_ptr.push_back(col_ptr.back() + count[col]);
    col_kmax.push_back(-1);
  }
  std::vector<int> col_sites(col_ptr.back());
  std::vector<int> fillpos(col_ids.size(), 0);
  for (size_t s = 0; s < S; ++s) {
    const int* b = &box[s * 6];
    if (b[0] > b[1]) continue;
    for (int i = b[0]; i <= b[1]; ++i)
      for (int j = b[2]; j <= b[3]; ++j) {
        const int sl = slot[(size_t)i * ny + j];
        col_sites[col_ptr[sl] + fillpos[sl]++] = (int)s;
        col_kmax[sl] = std::max(col_kmax[sl], b[5]);
      }
  }
  nb.ncols = col_ids.size();
  // far-field geometry: FP32 offsets of centres from the first centre of each tile
  std::vector<float> offx(nx), offy(ny), offz(nz);
  for (int i = 0; i < nx; ++i) offx[i] = (float)(xcen[i] - xcen[(i / kTile) * kTile]);
  for (int j = 0; j < ny; ++j) offy[j] = (float)(ycen[j] - ycen[(j / kTile) * kTile]);
  for (int k = 0; k < nz; ++k) offz[k] = (float)(zcen[k] - zcen[(k / kTile) * kTile]);
  hipStream_t st = ctx->stream;
  hipError_t e;
  if ((e = upload_vec(nb.box, box, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_ptr, col_ptr, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_sites, col_sites, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_ids, col_ids, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_kmax, col_kmax, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.xcen, xcen, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.ycen, ycen, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.zcen, zcen, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.offx, offx, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.offy, offy, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.offz, offz, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  nb.adj_table_valid = false;
  nb.valid = true;
  return hipSuccess;
}

hipError_t launch_forward_mixed(jgm_context* ctx, int kind, const double* model, double* out,
                                double* raw_out) {
  if (!ctx->near.valid) {
    hipError_t e = build_near_boxes(ctx);
    if (e != hipSuccess) return e;
  }
  const bool q = ctx->near.quad;
  switch (mask_of(kind)) {
    case kMaskScalar:
      return q ? forward_mixed_impl<kMaskScalar, true>(ctx, kind, model, out, raw_out)
               : forward_mixed_impl<kMaskScalar, false>(ctx, kind, model, out, raw_out);
    case kMaskAll:
      return q ? forward_mixed_impl<kMaskAll, true>(ctx, kind, model, out, raw_out)
               : forward_mixed_impl<kMaskAll, false>(ctx, kind, model, out, raw_out);
    default:
      return q ? forward_mixed_impl<kMaskTensor, true>(ctx, kind, model, out, raw_out)
               : forward_mixed_impl<kMaskTensor, false>(ctx, kind, model, out, raw_out);
  }
}

hipError_t launch_adjoint_mixed(jgm_context* ctx, int kind, const double* /*model*/,
                                const double* raw_misfit, double* grad) {
  if (!ctx->near.valid) {
    hipError_t e = build_near_boxes(ctx);
    if (e != hipSuccess) return e;
  }
  const bool q = ctx->near.quad;
  switch (mask_of(kind)) {
    case kMaskScalar:
      return q ? adjoint_mixed_impl<kMaskScalar, true>(ctx, kind, raw_misfit, grad)
               : adjoint_mixed_impl<kMaskScalar, false>(ctx, kind, raw_misfit, grad);
    case kMaskAll:
      return q ? adjoint_mixed_impl<kMaskAll, true>(ctx, kind, raw_misfit, grad)
               : adjoint_mixed_impl<kMaskAll, false>(ctx, kind, raw_misfit, grad);
    default:
      return q ? adjoint_mixed_impl<kMaskTensor, true>(ctx, kind, raw_misfit, grad)
               : adjoint_mixed_impl<kMaskTensor, false>(ctx, kind, raw_misfit, grad);
  }
}

}  // namespace jgm
