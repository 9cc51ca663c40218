// FP64 closed-form path: corner-first forward and adjoint, 1-D background,
// epilogues (component assembly, magnetic projection, data transforms) and the
// per-cell sensitivity rows.
//
// Layouts in HBM
//   corner coordinates Xc[nx+1], Yc[ny+1], Zc[nz+1]      (FP64)
//   corner weights     w[(I*(ny+1) + J)*(nz+1) + K]       (FP64, K fastest)
//   forward partials   partial[(chunk*NC + m)*S + s]      (FP64, sites fastest)
//   adjoint partials   hpart[chunk*Ncorner + q]           (FP64, corners fastest)
//   background table   bgT[(s*L + j)*7 + m]               (FP64, m = gz,xx,yy,zz,xy,xz,yz)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "context.hpp"
#include "prism_device.hpp"
#include "epilogue.hpp"

namespace jgm {

namespace {

constexpr int kBlock = 256;
// Waves per corner-kernel launch (~64 per SIMD over 256 CUs x 4 SIMDs): the sites or
// corners are cut into chunks to reach about this many; against 16384 the headline FP64
// step went from 88.1 to 85.5 ms.  The partial sums (~0.24 GB per model) are reduced in
// fixed order.
constexpr int kTargetWaves = 65536;

// w(I,J,K) = sum_{a,b,c} sigma(a,b,c) rho(I-a, J-b, K-c), rho zero outside the mesh.
// For a batch of nb models (model m at rho + m * model_ld) the weights of one corner
// are stored together: w[q * nb + m].
__global__ __launch_bounds__(kBlock) void k_corner_weights(const double* __restrict__ rho,
                                                           long model_ld, int nb, int nx, int ny,
                                                           int nz, double* __restrict__ w) {
  const long ncx = nx + 1, ncy = ny + 1, ncz = nz + 1;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncx * ncy * ncz) return;
  const int K = (int)(q % ncz);
  const long t = q / ncz;
  const int J = (int)(t % ncy);
  const int I = (int)(t / ncy);
  for (int m = 0; m < nb; ++m) {
    const double* r = rho + m * model_ld;
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int i = I - a, j = J - b, k = K - c;
          if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) continue;
          const double v = r[((long)i * ny + j) * nz + k];
          acc += ((a + b + c) & 1) ? v : -v;
        }
    w[q * nb + m] = acc;
  }
}

// Forward: one lane per site; the wave walks a chunk of corner columns (I, J)
// and all K.  Corner coordinates and weights are wave-uniform (scalar loads).  NB
// models share every corner evaluation (weights w[q * NB + b]); the partial sums of
// model b go to partial[((b * nchunks + chunk) * NC + m) * S + s].
template <int MASK, int NB>
__global__ __launch_bounds__(kBlock) void k_forward_corners(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const double* __restrict__ Xc, const double* __restrict__ Yc,
    const double* __restrict__ Zc, int ncy, int ncz, const double* __restrict__ w, long ncols,
    long cols_per_chunk, double* __restrict__ partial) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = s < S ? s : S - 1;
  const double px = sx[sl], py = sy[sl], pz = sz[sl];
  const long col0 = (long)blockIdx.y * cols_per_chunk;
  const long col1 = col0 + cols_per_chunk < ncols ? col0 + cols_per_chunk : ncols;
  double acc[NB][NC];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[b][m] = 0.0;
  for (long col = col0; col < col1; ++col) {
    const int I = (int)(col / ncy);
    const int J = (int)(col - (long)I * ncy);
    const double x = px - Xc[I];
    const double y = py - Yc[J];
    const double* wc = w + col * ncz * NB;
    for (int K = 0; K < ncz; ++K) {
      const double z = pz - Zc[K];
      double t[NC];
      corner_terms_f64<MASK>(x, y, z, t);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const double wk = wc[K * NB + b];
#pragma unroll
        for (int m = 0; m < NC; ++m) acc[b][m] = fma(wk, t[m], acc[b][m]);
      }
    }
  }
  if (s < S) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int m = 0; m < NC; ++m)
        partial[(((long)b * gridDim.y + blockIdx.y) * NC + m) * S + s] = acc[b][m];
  }
}

// Background table: per site and layer the 7 box-term values of the layer minus
// its part inside the mesh (jif3D/Gravity/GravityBackground.cpp:24-137).
__device__ inline double inf_sheet(double measz, double top, double bottom) {
  const double thick = bottom - top;
  if ((measz <= top) || (measz >= bottom))
    return 2.0 * kPi * kGravConst * thick * (top - measz >= 0 ? 1.0 : -1.0);
  return 2.0 * kPi * kGravConst * (thick - 2.0 * (measz - top));
}
__device__ inline double uzz_inf_sheet(double measz, double top, double bottom) {
  if ((measz <= top) || (measz >= bottom)) return 0.0;
  return 4.0 * kPi * kGravConst;
}

__global__ __launch_bounds__(kBlock) void k_background_table(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const double* __restrict__ thick, int L, double x0, double y0, double z0,
    double xwidth, double ywidth, double zwidth, double* __restrict__ bgT) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const double px = sx[s], py = sy[s], pz = sz[s];
  double currtop = z0;
  for (int j = 0; j < L; ++j) {
    const double th = thick[j];
    const double currbottom = currtop + th;
    double v[7];
    v[0] = inf_sheet(pz, currtop, currbottom);
    for (int m = 1; m < 7; ++m) v[m] = 0.0;
    v[3] = uzz_inf_sheet(pz, currtop, currbottom);
    double boxh = -1.0;
    if (currtop < zwidth && currbottom <= zwidth) boxh = th;
    if (currtop < zwidth && currbottom > zwidth) boxh = zwidth - currtop;
    if (boxh >= 0.0) {
      double g[1], t[6];
      box_terms_f64<kMaskScalar>(px, py, pz, x0, y0, currtop, xwidth, ywidth, boxh, g);
      box_terms_f64<kMaskTensor>(px, py, pz, x0, y0, currtop, xwidth, ywidth, boxh, t);
      v[0] -= g[0];
      for (int m = 0; m < 6; ++m) v[1 + m] -= t[m];
    }
    double* o = bgT + ((long)s * L + j) * 7;
    for (int m = 0; m < 7; ++m) o[m] = v[m];
    currtop += th;
  }
}

// Sum the chunk partials, scale, add the background and assemble the raw data of
// one site; then apply the transform.  u[] holds gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_forward_finish(
    const double* __restrict__ partial, int nchunks, int S, const double* __restrict__ bgT,
    const double* __restrict__ bgvals, int L, int kind, FieldDir fd, int tf, double tfbx,
    double tfby, double tfbz, double* __restrict__ out, double* __restrict__ raw_out) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double acc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = 0.0;
  for (int c = 0; c < nchunks; ++c)
#pragma unroll
    for (int m = 0; m < NC; ++m) acc[m] += partial[((long)c * NC + m) * S + s];
  double sc[NC];
  comp_scales<MASK>(sc);
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] *= sc[m];
  double u[7];
  spread_components<MASK>(acc, u);
  finish_site(u, s, bgT, bgvals, L, kind, fd, tf, tfbx, tfby, tfbz, out, raw_out);
}

// Misfit of transformed data -> misfit of raw data: D^T r per site
// (FullSensitivityGravMagCalculator.h:184-208; ThreeDGravMagImplementation.h:212-219).
__global__ __launch_bounds__(kBlock) void k_misfit_to_raw(const double* __restrict__ misfit,
                                                          const double* __restrict__ raw, int S,
                                                          int kraw, int tf, double tfbx,
                                                          double tfby, double tfbz,
                                                          double* __restrict__ rm) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  if (tf == JGM_TF_NONE) {
    for (int m = 0; m < kraw; ++m) rm[(long)s * kraw + m] = misfit[(long)s * kraw + m];
    return;
  }
  const double r = misfit[s];
  const double* in = raw + (long)s * kraw;
  double d[9];
  if (tf == JGM_TF_FTG_INVARIANT) {  // GravityTransforms.h:80-91
    d[0] = in[4] + in[8]; d[1] = -in[3]; d[2] = -in[6];
    d[3] = -in[1]; d[4] = in[0] + in[8]; d[5] = -in[7];
    d[6] = -in[2]; d[7] = -in[5]; d[8] = in[0] + in[4];
  } else if (tf == JGM_TF_TOTAL_FIELD) {  // MagneticTransforms.h:66-72
    const double t = sqrt(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    d[0] = in[0] / t; d[1] = in[1] / t; d[2] = in[2] / t;
  } else {
    d[0] = tfbx; d[1] = tfby; d[2] = tfbz;
  }
  for (int m = 0; m < kraw; ++m) rm[(long)s * kraw + m] = d[m] * r;
}

// Per-site coefficients of the 7 corner sums for the adjoint:
// coef[s][m] = sum_k rraw[s][k] * dJ_k/dS_m where S_m is the corner sum of component m
// (including its sign and G).  Stored in the kernel's component order for MASK at
// coef[s * ld + m] (ld = NC for one model; NB * NC with model b at coef + b * NC).
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_adjoint_coef(const double* __restrict__ rm, int S,
                                                         int kind, FieldDir fd,
                                                         double* __restrict__ coef, int ld) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double c[7];  // gz, xx, yy, zz, xy, xz, yz (before sign*G)
  site_coefficients(rm, s, kind, fd, c);
  double sc[NC];
  comp_scales<MASK>(sc);
  int k = 0;
  if (MASK & kGz) coef[(long)s * ld + k] = c[0] * sc[k], ++k;
  if (MASK & kDiag)
    for (int m = 1; m <= 3; ++m) coef[(long)s * ld + k] = c[m] * sc[k], ++k;
  if (MASK & kOff)
    for (int m = 4; m <= 6; ++m) coef[(long)s * ld + k] = c[m] * sc[k], ++k;
}

// Adjoint: one lane per corner; the wave walks a chunk of sites (wave-uniform).  NB
// residual sets share every corner evaluation (coefficients coef[(s * NB + b) * NC + m]);
// the corner sums of set b go to hpart[(b * nchunks + chunk) * ncorner + q].
template <int MASK, int NB>
__global__ __launch_bounds__(kBlock) void k_adjoint_corners(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const double* __restrict__ coef, int S, int sites_per_chunk,
    const double* __restrict__ Xc, const double* __restrict__ Yc,
    const double* __restrict__ Zc, int ncy, int ncz, long ncorner, double* __restrict__ hpart) {
  constexpr int NC = Ncomp<MASK>::value;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ql = q < ncorner ? q : ncorner - 1;
  const int K = (int)(ql % ncz);
  const long t = ql / ncz;
  const int J = (int)(t % ncy);
  const int I = (int)(t / ncy);
  const double cx = Xc[I], cy = Yc[J], cz = Zc[K];
  const int s0 = blockIdx.y * sites_per_chunk;
  const int s1 = s0 + sites_per_chunk < S ? s0 + sites_per_chunk : S;
  double h[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) h[b] = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double x = sx[s] - cx, y = sy[s] - cy, z = sz[s] - cz;
    double tv[NC];
    corner_terms_f64<MASK>(x, y, z, tv);
    const double* cs = coef + (long)s * NB * NC;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int m = 0; m < NC; ++m) h[b] = fma(cs[b * NC + m], tv[m], h[b]);
  }
  if (q < ncorner) {
#pragma unroll
    for (int b = 0; b < NB; ++b) hpart[((long)b * gridDim.y + blockIdx.y) * ncorner + q] = h[b];
  }
}

__global__ __launch_bounds__(kBlock) void k_sum_chunks(const double* __restrict__ part,
                                                       int nchunks, long n,
                                                       double* __restrict__ out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double a = 0.0;
  for (int c = 0; c < nchunks; ++c) a += part[(long)c * n + q];
  out[q] = a;
}

// First pass of the chunk reduction: group g sums chunks [g per, (g + 1) per) of every
// row; rows are coalesced across lanes.
__global__ __launch_bounds__(kBlock) void k_sum_chunk_groups(const double* __restrict__ part,
                                                             int nchunks, int per, long n,
                                                             double* __restrict__ grp) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int c0 = (int)blockIdx.y * per;
  const int c1 = c0 + per < nchunks ? c0 + per : nchunks;
  double a = 0.0;
  for (int c = c0; c < c1; ++c) a += part[(long)c * n + q];
  grp[(long)blockIdx.y * n + q] = a;
}

// g_c = 2 sum_{a,b,c} sigma(a,b,c) H(i+a, j+b, k+c)
__global__ __launch_bounds__(kBlock) void k_corner_difference(const double* __restrict__ H,
                                                              int nx, int ny, int nz,
                                                              double* __restrict__ grad) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)nx * ny * nz;
  if (c >= n) return;
  const int k = (int)(c % nz);
  const long t = c / nz;
  const int j = (int)(t % ny);
  const int i = (int)(t / ny);
  const long ncy = ny + 1, ncz = nz + 1;
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const double v = H[((long)(i + a) * ncy + (j + b)) * ncz + (k + cc)];
        acc += ((a + b + cc) & 1) ? v : -v;
      }
  grad[c] = 2.0 * acc;
}

// Background gradient: g[N + j] = 2 sum_s sum_k rraw[s][k] * Jbg[s][k][j].
__global__ __launch_bounds__(kBlock) void k_background_gradient(
    const double* __restrict__ rm, const double* __restrict__ bgT, int S, int L, int kind,
    FieldDir fd, double* __restrict__ gbg) {
  const int j = blockIdx.x;
  __shared__ double red[kBlock];
  double acc = 0.0;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const double* b = bgT + ((long)s * L + j) * 7;
    if (kind == JGM_KIND_SCALAR) {
      acc += rm[s] * b[0];
    } else if (kind == JGM_KIND_TENSOR || kind == JGM_KIND_SCALAR_TENSOR) {
      const int k = kind == JGM_KIND_SCALAR_TENSOR ? 10 : 9;
      const double* r = rm + (long)s * k;
      int o = 0;
      if (kind == JGM_KIND_SCALAR_TENSOR) acc += r[o++] * b[0];
      acc += r[o + 0] * b[1] + r[o + 1] * b[4] + r[o + 2] * b[5] + r[o + 3] * b[4] +
             r[o + 4] * b[2] + r[o + 5] * b[6] + r[o + 6] * b[5] + r[o + 7] * b[6] +
             r[o + 8] * b[3];
    } else {
      const double* r = rm + (long)s * 3;
      const double Bx = (b[1] * fd.bx + b[4] * fd.by + b[5] * fd.bz) * fd.scale;
      const double By = (b[4] * fd.bx + b[2] * fd.by + b[6] * fd.bz) * fd.scale;
      const double Bz = (b[5] * fd.bx + b[6] * fd.by + b[3] * fd.bz) * fd.scale;
      acc += r[0] * Bx + r[1] * By + r[2] * Bz;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) gbg[j] = 2.0 * red[0];
}

// Raw sensitivity rows for a range of sites (reference layout), one lane per
// parameter column, per-cell 8-corner evaluation like the OMP imps.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_sensitivity_rows(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int site0, const double* __restrict__ xlo, const double* __restrict__ ylo,
    const double* __restrict__ zlo, const double* __restrict__ xsz,
    const double* __restrict__ ysz, const double* __restrict__ zsz, int nx, int ny, int nz,
    const double* __restrict__ bgT, int L, int kind, FieldDir fd, double mscale,
    double* __restrict__ rows) {
  constexpr int NC = Ncomp<MASK>::value;
  const long N = (long)nx * ny * nz;
  const long np = kind == JGM_KIND_MAGNETIZATION ? 3 * N : N + L;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= np) return;
  const int sl = blockIdx.y;  // local site index
  const int s = site0 + sl;
  if (kind == JGM_KIND_MAGNETIC_GRADIENT || kind == JGM_KIND_MAGNETIZATION) {
    // OMPMagneticGradImp.cpp:13-106 (no background columns: zeros) and
    // OMPMagnetizationImp.cpp:25-117 (column p N + cell of component p)
    double* base = rows + (long)sl * raw_per_site(kind) * np + c;
    const int p = kind == JGM_KIND_MAGNETIZATION ? (int)(c / N) : 0;
    const long cell = c - p * N;
    if (cell >= N) {
      base[0] = 0.0;
      return;
    }
    const int k = (int)(cell % nz);
    const long t = cell / nz;
    const int j = (int)(t % ny);
    const int i = (int)(t / ny);
    double v[NC], u[7], w[7];
    box_terms_f64<MASK>(sx[s], sy[s], sz[s], xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
    spread_components<MASK>(v, u);  // gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz
    if (kind == JGM_KIND_MAGNETIC_GRADIENT) {
      box_terms_f64<MASK>(sx[s], sy[s], sz[s] + 1.0, xlo[i], ylo[j], zlo[k], xsz[i], ysz[j],
                          zsz[k], v);
      spread_components<MASK>(v, w);
      base[0] = (u[5] * fd.bx + u[6] * fd.by + u[3] * fd.bz) * mscale -
                (w[5] * fd.bx + w[6] * fd.by + w[3] * fd.bz) * mscale;
    } else {  // column p of T = [[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]]
      base[0] = (p == 0 ? u[1] : p == 1 ? u[4] : u[5]) * mscale;
      base[np] = (p == 0 ? u[4] : p == 1 ? u[2] : u[6]) * mscale;
      base[2 * np] = (p == 0 ? u[5] : p == 1 ? u[6] : u[3]) * mscale;
    }
    return;
  }
  double u[7] = {0, 0, 0, 0, 0, 0, 0};
  if (c < N) {
    const int k = (int)(c % nz);
    const long t = c / nz;
    const int j = (int)(t % ny);
    const int i = (int)(t / ny);
    double v[NC];
    box_terms_f64<MASK>(sx[s], sy[s], sz[s], xlo[i], ylo[j], zlo[k], xsz[i], ysz[j], zsz[k], v);
    int m = 0;
    if (MASK & kGz) u[0] = v[m++];
    if (MASK & kDiag) { u[1] = v[m++]; u[2] = v[m++]; u[3] = v[m++]; }
    if (MASK & kOff) { u[4] = v[m++]; u[5] = v[m++]; u[6] = v[m++]; }
  } else {
    const double* b = bgT + ((long)s * L + (c - N)) * 7;
    for (int m = 0; m < 7; ++m) u[m] = b[m];
  }
  const int kraw = (int)raw_per_site(kind);
  double* base = rows + (long)sl * kraw * np + c;
  if (kind == JGM_KIND_SCALAR) {
    base[0] = u[0];
  } else if (kind == JGM_KIND_TENSOR || kind == JGM_KIND_SCALAR_TENSOR) {
    int o = 0;
    if (kind == JGM_KIND_SCALAR_TENSOR) base[(long)(o++) * np] = u[0];
    const double t9[9] = {u[1], u[4], u[5], u[4], u[2], u[6], u[5], u[6], u[3]};
    for (int m = 0; m < 9; ++m) base[(long)(o + m) * np] = t9[m];
  } else {
    base[0] = (u[1] * fd.bx + u[4] * fd.by + u[5] * fd.bz) * fd.scale;
    base[np] = (u[4] * fd.bx + u[2] * fd.by + u[6] * fd.bz) * fd.scale;
    base[2 * np] = (u[5] * fd.bx + u[6] * fd.by + u[3] * fd.bz) * fd.scale;
  }
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

int mask_of(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return kMaskScalar;
    case JGM_KIND_SCALAR_TENSOR: return kMaskAll;
    default: return kMaskTensor;
  }
}

int ncomp_of(int mask) {
  return ((mask & kGz) ? 1 : 0) + ((mask & kDiag) ? 3 : 0) + ((mask & kOff) ? 3 : 0);
}

// Launch k_forward_corners / k_adjoint_corners for a batch size known at run time.
template <int MASK, class... Args>
void launch_forward_corners(int nb, dim3 grid, hipStream_t st, Args... args) {
  switch (nb) {
    case 8: hipLaunchKernelGGL((k_forward_corners<MASK, 8>), grid, dim3(kBlock), 0, st, args...); break;
    case 4: hipLaunchKernelGGL((k_forward_corners<MASK, 4>), grid, dim3(kBlock), 0, st, args...); break;
    case 2: hipLaunchKernelGGL((k_forward_corners<MASK, 2>), grid, dim3(kBlock), 0, st, args...); break;
    default: hipLaunchKernelGGL((k_forward_corners<MASK, 1>), grid, dim3(kBlock), 0, st, args...); break;
  }
}
template <int MASK, class... Args>
void launch_adjoint_corners(int nb, dim3 grid, hipStream_t st, Args... args) {
  switch (nb) {
    case 8: hipLaunchKernelGGL((k_adjoint_corners<MASK, 8>), grid, dim3(kBlock), 0, st, args...); break;
    case 4: hipLaunchKernelGGL((k_adjoint_corners<MASK, 4>), grid, dim3(kBlock), 0, st, args...); break;
    case 2: hipLaunchKernelGGL((k_adjoint_corners<MASK, 2>), grid, dim3(kBlock), 0, st, args...); break;
    default: hipLaunchKernelGGL((k_adjoint_corners<MASK, 1>), grid, dim3(kBlock), 0, st, args...); break;
  }
}

// Forward of nb in {1, 2, 4, 8} models (model b at models + b * model_ld, its data at
// out + b * out_ld, raw data at raw_out + b * raw_ld when raw_out is given).
template <int MASK>
hipError_t forward_impl(jgm_context* ctx, int kind, int nb, const double* models, size_t model_ld,
                        double* out, size_t out_ld, double* raw_out, size_t raw_ld) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const int ncx = (int)ctx->nx + 1, ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long ncols = (long)ncx * ncy;
  const int site_blocks = (S + kBlock - 1) / kBlock;
  int nchunks = 1;
  if (N > 0) {
    const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
    long want = (kTargetWaves + waves_per_chunk - 1) / waves_per_chunk;
    want = std::min<long>(want, ncols);
    nchunks = (int)std::max<long>(1, want);
  }
  const long cols_per_chunk = (ncols + nchunks - 1) / nchunks;
  nchunks = (int)((ncols + cols_per_chunk - 1) / cols_per_chunk);
  hipError_t e;
  if (N > 0) {
    e = ctx->weights.ensure(sizeof(double) * (size_t)nb * ncols * ncz);
    if (e != hipSuccess) return e;
    e = ctx->partial.ensure(sizeof(double) * (size_t)nb * nchunks * NC * S);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_corner_weights, dim3(blocks_for(ncols * ncz)), dim3(kBlock), 0, st,
                       models, (long)model_ld, nb, (int)ctx->nx, (int)ctx->ny, (int)ctx->nz,
                       ctx->weights.as<double>());
    ProfScope ps(ctx, MASK == kMaskScalar ? "forward_corners_fp64_gz"
                      : MASK == kMaskAll  ? "forward_corners_fp64_gz_ftg"
                                          : "forward_corners_fp64_ftg");
    launch_forward_corners<MASK>(nb, dim3(site_blocks, nchunks), st,
                                 ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                                 ctx->d_sz.as<const double>(), S, ctx->d_Xc.as<const double>(),
                                 ctx->d_Yc.as<const double>(), ctx->d_Zc.as<const double>(), ncy,
                                 ncz, ctx->weights.as<const double>(), ncols, cols_per_chunk,
                                 ctx->partial.as<double>());
  } else {
    nchunks = 0;
  }
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  if (nchunks > 0 && (e = ctx->partsum.ensure(sizeof(double) * (size_t)NC * S)) != hipSuccess)
    return e;
  for (int b = 0; b < nb; ++b) {
    if (nchunks > 0 &&
        (e = launch_reduce_chunks(ctx, ctx->partial.as<const double>() + (size_t)b * nchunks * NC * S,
                                  nchunks, (long)NC * S, ctx->partsum.as<double>())) != hipSuccess)
      return e;
    hipLaunchKernelGGL((k_forward_finish<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st,
                       ctx->partsum.as<const double>(), nchunks > 0 ? 1 : 0, S,
                       ctx->bgT.as<const double>(), models + b * model_ld + N, L, kind,
                       field_dir(ctx), ctx->transform, tf_bx, tf_by, tf_bz, out + b * out_ld,
                       raw_out ? raw_out + b * raw_ld : nullptr);
  }
  return hipGetLastError();
}

// Adjoint of nb in {1, 2, 4, 8} raw residual sets (set b at rm + b * rm_ld, its
// gradient at grad + b * grad_ld).
template <int MASK>
hipError_t adjoint_impl(jgm_context* ctx, int kind, int nb, const double* rm, size_t rm_ld,
                        double* grad, size_t grad_ld) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long nq = (long)ncorners(ctx);
  hipError_t e = ctx->coef.ensure(sizeof(double) * (size_t)nb * S * NC);
  if (e != hipSuccess) return e;
  for (int b = 0; b < nb; ++b)
    hipLaunchKernelGGL((k_adjoint_coef<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st,
                       rm + b * rm_ld, S, kind, field_dir(ctx), ctx->coef.as<double>() + b * NC,
                       nb * NC);
  if (N > 0) {
    const long corner_waves = (nq + 63) / 64;
    int nchunks = (int)std::min<long>(S, std::max<long>(1, (kTargetWaves + corner_waves - 1) / corner_waves));
    const int per = (S + nchunks - 1) / nchunks;
    nchunks = (S + per - 1) / per;
    e = ctx->hpart.ensure(sizeof(double) * (size_t)nb * nchunks * nq);
    if (e != hipSuccess) return e;
    {
    ProfScope ps(ctx, MASK == kMaskScalar ? "adjoint_corners_fp64_gz"
                      : MASK == kMaskAll  ? "adjoint_corners_fp64_gz_ftg"
                                          : "adjoint_corners_fp64_ftg");
    launch_adjoint_corners<MASK>(nb, dim3(blocks_for(nq), nchunks), st,
                                 ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                                 ctx->d_sz.as<const double>(), ctx->coef.as<const double>(), S,
                                 per, ctx->d_Xc.as<const double>(), ctx->d_Yc.as<const double>(),
                                 ctx->d_Zc.as<const double>(), ncy, ncz, nq,
                                 ctx->hpart.as<double>());
    }
    if (nchunks > 1) {
      e = ctx->hsum.ensure(sizeof(double) * (size_t)nq);
      if (e != hipSuccess) return e;
    }
    for (int b = 0; b < nb; ++b) {
      const double* H = ctx->hpart.as<const double>() + (size_t)b * nchunks * nq;
      if (nchunks > 1) {
        hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(nq)), dim3(kBlock), 0, st, H, nchunks,
                           nq, ctx->hsum.as<double>());
        H = ctx->hsum.as<const double>();
      }
      hipLaunchKernelGGL(k_corner_difference, dim3(blocks_for(N)), dim3(kBlock), 0, st, H,
                         (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, grad + b * grad_ld);
    }
  }
  if (L > 0) {
    for (int b = 0; b < nb; ++b) {
      hipError_t eb = launch_background_gradient(ctx, kind, rm + b * rm_ld, grad + b * grad_ld + N);
      if (eb != hipSuccess) return eb;
    }
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Magnetic vertical gradient (OMPMagneticGradImp.cpp:13-106) and magnetization
// (OMPMagnetizationImp.cpp:25-117).  The vertical gradient is the difference of the
// susceptibility Bz at the site and 1 m below it, so both positions run through the
// tensor corner kernels as 2S virtual sites.  The magnetization model has three values
// per cell, B = 1e9 mu0/(4 pi G) T M, evaluated corner-first with three corner weight
// sets in one pass.
// ---------------------------------------------------------------------------

// Virtual sites: (x, y, z) for s < S and (x, y, z + dz) for s >= S.
__global__ __launch_bounds__(kBlock) void k_pair_sites(const double* __restrict__ sx,
                                                       const double* __restrict__ sy,
                                                       const double* __restrict__ sz, int S,
                                                       double dz, double* __restrict__ vs) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 2 * S) return;
  const int s = v < S ? v : v - S;
  vs[v] = sx[s];
  vs[2 * S + v] = sy[s];
  vs[4 * S + v] = sz[s] + (v < S ? 0.0 : dz);
}

// Bz projection of the tensor kernel's components (xx, yy, zz, xy, xz, yz):
// (T b)_z = Uxz bx + Uyz by + Uzz bz, times F mu0/(4 pi) / G = F 1e-7 / G.
// dst[i] += src[i]
__global__ __launch_bounds__(kBlock) void k_add_f64(double* __restrict__ dst,
                                                    const double* __restrict__ src, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

// Site s of the list goes to row idx[s] of the output (row s without idx).
__global__ __launch_bounds__(kBlock) void k_maggrad_finish(const double* __restrict__ partial,
                                                           int nchunks, int S, FieldDir fd,
                                                           double fs, const int* __restrict__ idx,
                                                           double* __restrict__ out,
                                                           double* __restrict__ raw_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int S2 = 2 * S;
  double bz[2];
  for (int h = 0; h < 2; ++h) {
    const int v = s + h * S;
    double zz = 0.0, xz = 0.0, yz = 0.0;
    for (int c = 0; c < nchunks; ++c) {
      zz += partial[((long)c * 6 + 2) * S2 + v];
      xz += partial[((long)c * 6 + 4) * S2 + v];
      yz += partial[((long)c * 6 + 5) * S2 + v];
    }
    // tensor signs: diagonal +G, off-diagonal -G (BasicGravElements.cpp:151, :201)
    bz[h] = (-kGravConst * xz * fd.bx - kGravConst * yz * fd.by + kGravConst * zz * fd.bz) *
            fs * (1e-7 / kGravConst);
  }
  const double v = bz[0] - bz[1];
  const int o = idx ? idx[s] : s;
  out[o] = v;
  if (raw_out) raw_out[o] = v;
}

// Adjoint coefficients of the 2S virtual sites: +r at the site, -r 1 m below.
// Misfit of listed site s at rm[idx[s]] (rm[s] without idx).
__global__ __launch_bounds__(kBlock) void k_maggrad_coef(const double* __restrict__ rm, int S,
                                                         FieldDir fd, double fs,
                                                         const int* __restrict__ idx,
                                                         double* __restrict__ coef) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 2 * S) return;
  const int s = v < S ? v : v - S;
  const double rs = rm[idx ? idx[s] : s];
  const double r = (v < S ? rs : -rs) * fs * (1e-7 / kGravConst);
  double* c = coef + (long)v * 6;  // xx, yy, zz, xy, xz, yz with their sign * G
  c[0] = 0.0;
  c[1] = 0.0;
  c[2] = r * fd.bz * kGravConst;
  c[3] = 0.0;
  c[4] = -r * fd.bx * kGravConst;
  c[5] = -r * fd.by * kGravConst;
}

// B_i = sum_j T_ij w_j over corners, T_ij = sign_ij F_ij (G and the unit factor
// applied in the finish); w = three corner weight sets [3][ncols * ncz].
__global__ __launch_bounds__(kBlock) void k_forward_corners_mag3(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const double* __restrict__ Xc, const double* __restrict__ Yc,
    const double* __restrict__ Zc, int ncy, int ncz, const double* __restrict__ w, long nq,
    long ncols, long cols_per_chunk, double* __restrict__ partial) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sl = s < S ? s : S - 1;
  const double px = sx[sl], py = sy[sl], pz = sz[sl];
  const long col0 = (long)blockIdx.y * cols_per_chunk;
  const long col1 = col0 + cols_per_chunk < ncols ? col0 + cols_per_chunk : ncols;
  double bx = 0.0, by = 0.0, bz = 0.0;
  for (long col = col0; col < col1; ++col) {
    const int I = (int)(col / ncy);
    const int J = (int)(col - (long)I * ncy);
    const double x = px - Xc[I];
    const double y = py - Yc[J];
    const double* wc = w + col * ncz;
    for (int K = 0; K < ncz; ++K) {
      const double z = pz - Zc[K];
      const double wx = wc[K], wy = wc[nq + K], wz = wc[2 * nq + K];
      double t[6];  // xx, yy, zz, xy, xz, yz
      corner_terms_f64<kMaskTensor>(x, y, z, t);
      bx += t[0] * wx - t[3] * wy - t[4] * wz;
      by += -t[3] * wx + t[1] * wy - t[5] * wz;
      bz += -t[4] * wx - t[5] * wy + t[2] * wz;
    }
  }
  if (s < S) {
    partial[((long)blockIdx.y * 3 + 0) * S + s] = bx;
    partial[((long)blockIdx.y * 3 + 1) * S + s] = by;
    partial[((long)blockIdx.y * 3 + 2) * S + s] = bz;
  }
}

__global__ __launch_bounds__(kBlock) void k_mag3_finish(
    const double* __restrict__ partial, int nchunks, int S, double scale, int tf, double tfbx,
    double tfby, double tfbz, double* __restrict__ out, double* __restrict__ raw_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double raw[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < nchunks; ++c)
    for (int i = 0; i < 3; ++i) raw[i] += partial[((long)c * 3 + i) * S + s];
  for (int i = 0; i < 3; ++i) raw[i] *= scale;
  store_site(raw, 3, s, tf, tfbx, tfby, tfbz, out, raw_out);
}

// Adjoint corner sums of the three model components: H_j = sum_s sum_i c_i(s) T_ij.
__global__ __launch_bounds__(kBlock) void k_adjoint_corners_mag3(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const double* __restrict__ coef, int S, int sites_per_chunk,
    const double* __restrict__ Xc, const double* __restrict__ Yc,
    const double* __restrict__ Zc, int ncy, int ncz, long ncorner, double* __restrict__ hpart) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ql = q < ncorner ? q : ncorner - 1;
  const int K = (int)(ql % ncz);
  const long t = ql / ncz;
  const int J = (int)(t % ncy);
  const int I = (int)(t / ncy);
  const double cx = Xc[I], cy = Yc[J], cz = Zc[K];
  const int s0 = blockIdx.y * sites_per_chunk;
  const int s1 = s0 + sites_per_chunk < S ? s0 + sites_per_chunk : S;
  double hx = 0.0, hy = 0.0, hz = 0.0;
  for (int s = s0; s < s1; ++s) {
    double tv[6];
    corner_terms_f64<kMaskTensor>(sx[s] - cx, sy[s] - cy, sz[s] - cz, tv);
    const double* c = coef + (long)s * 3;
    hx += c[0] * tv[0] - c[1] * tv[3] - c[2] * tv[4];
    hy += -c[0] * tv[3] + c[1] * tv[1] - c[2] * tv[5];
    hz += -c[0] * tv[4] - c[1] * tv[5] + c[2] * tv[2];
  }
  if (q < ncorner) {
    double* o = hpart + (long)blockIdx.y * 3 * ncorner + q;
    o[0] = hx;
    o[ncorner] = hy;
    o[2 * ncorner] = hz;
  }
}

__global__ __launch_bounds__(kBlock) void k_mag3_coef(const double* __restrict__ rm, int S,
                                                      double scale, double* __restrict__ coef) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3L * S) coef[i] = rm[i] * scale;
}

// chunking of a lane-per-site corner walk (as forward_impl)
void site_chunks(long S, long ncols, long* nchunks, long* cols_per_chunk) {
  const long site_blocks = (S + kBlock - 1) / kBlock;
  const long waves_per_chunk = site_blocks * (kBlock / 64);
  long want = (kTargetWaves + waves_per_chunk - 1) / waves_per_chunk;
  want = std::max<long>(1, std::min<long>(want, ncols));
  *cols_per_chunk = (ncols + want - 1) / want;
  *nchunks = (ncols + *cols_per_chunk - 1) / *cols_per_chunk;
}

// chunking of a lane-per-corner walk over sites (as adjoint_impl)
void corner_chunks(long S, long nq, int* nchunks, int* per) {
  const long corner_waves = (nq + 63) / 64;
  int n = (int)std::min<long>(S, std::max<long>(1, (kTargetWaves + corner_waves - 1) / corner_waves));
  *per = (int)((S + n - 1) / n);
  *nchunks = (int)((S + *per - 1) / *per);
}

double magnetization_scale() {
  const double mu0 = 4e-7 * kPi;  // OMPMagnetizationImp.cpp:35
  return 1.0e9 * mu0 / (4 * kPi * kGravConst) * kGravConst;  // factor times the G of T
}

// Vertical gradient at S listed sites (sx, sy, sz), written to rows idx[s] of out and
// raw_out (rows s without idx): the tensor corner sums at the site and 1 m below.
hipError_t maggrad_forward_at(jgm_context* ctx, const double* model, int S, const double* sx,
                              const double* sy, const double* sz, const int* idx, double* out,
                              double* raw_out) {
  hipStream_t st = ctx->stream;
  const int S2 = 2 * S;
  const long N = (long)ncells(ctx);
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long ncols = (long)((int)ctx->nx + 1) * ncy;
  long nchunks = 0, per = 1;
  hipError_t e;
  if ((e = ctx->vsites.ensure(sizeof(double) * 6 * (size_t)S)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_pair_sites, dim3(blocks_for(S2)), dim3(kBlock), 0, st, sx, sy, sz, S, 1.0,
                     ctx->vsites.as<double>());
  if (N > 0) {
    site_chunks(S2, ncols, &nchunks, &per);
    if ((e = ctx->weights.ensure(sizeof(double) * (size_t)ncols * ncz)) != hipSuccess) return e;
    if ((e = ctx->partial.ensure(sizeof(double) * (size_t)nchunks * 6 * S2)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_corner_weights, dim3(blocks_for(ncols * ncz)), dim3(kBlock), 0, st,
                       model, 0L, 1, (int)ctx->nx, (int)ctx->ny, (int)ctx->nz,
                       ctx->weights.as<double>());
    const double* vs = ctx->vsites.as<const double>();
    ProfScope ps(ctx, "forward_corners_fp64_maggrad");
    hipLaunchKernelGGL((k_forward_corners<kMaskTensor, 1>), dim3((S2 + kBlock - 1) / kBlock, nchunks),
                       dim3(kBlock), 0, st, vs, vs + S2, vs + 2 * S2, S2, ctx->d_Xc.as<double>(),
                       ctx->d_Yc.as<double>(), ctx->d_Zc.as<double>(), ncy, ncz,
                       ctx->weights.as<const double>(), ncols, per, ctx->partial.as<double>());
  }
  if (nchunks > 0) {
    if ((e = ctx->partsum.ensure(sizeof(double) * 6 * (size_t)S2)) != hipSuccess) return e;
    if ((e = launch_reduce_chunks(ctx, ctx->partial.as<const double>(), (int)nchunks, 6L * S2,
                                  ctx->partsum.as<double>())) != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL(k_maggrad_finish, dim3(blocks_for(S)), dim3(kBlock), 0, st,
                     ctx->partsum.as<const double>(), nchunks > 0 ? 1 : 0, S, field_dir(ctx),
                     ctx->fieldstrength, idx, out, raw_out);
  return hipGetLastError();
}

hipError_t maggrad_forward(jgm_context* ctx, const double* model, double* out,
                           double* raw_out) {
  return maggrad_forward_at(ctx, model, (int)ctx->S, ctx->d_sx.as<const double>(),
                            ctx->d_sy.as<const double>(), ctx->d_sz.as<const double>(), nullptr,
                            out, raw_out);
}

// Gradient of the vertical gradient's misfit at S listed sites (misfit of site s at
// rm[idx[s]]): stored into grad[0, N + L) (background columns zero, as
// OMPMagneticGradImp::CalcBackground), or added to grad[0, N) when add is set.
hipError_t maggrad_adjoint_at(jgm_context* ctx, int S, const double* sx, const double* sy,
                              const double* sz, const int* idx, const double* rm, double* grad,
                              bool add) {
  hipStream_t st = ctx->stream;
  const int S2 = 2 * S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long nq = (long)ncorners(ctx);
  hipError_t e;
  if (L > 0 && !add) {
    if ((e = launch_fill_zero(ctx, grad + N, (size_t)L)) != hipSuccess) return e;
  }
  if (N == 0 || S == 0) return hipSuccess;
  if ((e = ctx->vsites.ensure(sizeof(double) * 6 * (size_t)S)) != hipSuccess) return e;
  if ((e = ctx->coef.ensure(sizeof(double) * 6 * (size_t)S2)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_pair_sites, dim3(blocks_for(S2)), dim3(kBlock), 0, st, sx, sy, sz, S, 1.0,
                     ctx->vsites.as<double>());
  hipLaunchKernelGGL(k_maggrad_coef, dim3(blocks_for(S2)), dim3(kBlock), 0, st, rm, S,
                     field_dir(ctx), ctx->fieldstrength, idx, ctx->coef.as<double>());
  int nchunks, per;
  corner_chunks(S2, nq, &nchunks, &per);
  if ((e = ctx->hpart.ensure(sizeof(double) * (size_t)nchunks * nq)) != hipSuccess) return e;
  const double* vs = ctx->vsites.as<const double>();
  {
    ProfScope ps(ctx, "adjoint_corners_fp64_maggrad");
    hipLaunchKernelGGL((k_adjoint_corners<kMaskTensor, 1>), dim3(blocks_for(nq), nchunks),
                       dim3(kBlock), 0, st, vs, vs + S2, vs + 2 * S2, ctx->coef.as<const double>(),
                       S2, per, ctx->d_Xc.as<double>(), ctx->d_Yc.as<double>(),
                       ctx->d_Zc.as<double>(), ncy, ncz, nq, ctx->hpart.as<double>());
  }
  const double* H = ctx->hpart.as<const double>();
  if (nchunks > 1) {
    if ((e = ctx->hsum.ensure(sizeof(double) * (size_t)nq)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(nq)), dim3(kBlock), 0, st, H, nchunks, nq,
                       ctx->hsum.as<double>());
    H = ctx->hsum.as<const double>();
  }
  if (!add) {
    hipLaunchKernelGGL(k_corner_difference, dim3(blocks_for(N)), dim3(kBlock), 0, st, H,
                       (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, grad);
    return hipGetLastError();
  }
  if ((e = ctx->rt_tmp.ensure(sizeof(double) * (size_t)N)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_corner_difference, dim3(blocks_for(N)), dim3(kBlock), 0, st, H,
                     (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, ctx->rt_tmp.as<double>());
  hipLaunchKernelGGL(k_add_f64, dim3(blocks_for(N)), dim3(kBlock), 0, st, grad,
                     ctx->rt_tmp.as<const double>(), N);
  return hipGetLastError();
}

hipError_t maggrad_adjoint(jgm_context* ctx, const double* rm, double* grad) {
  return maggrad_adjoint_at(ctx, (int)ctx->S, ctx->d_sx.as<const double>(),
                            ctx->d_sy.as<const double>(), ctx->d_sz.as<const double>(), nullptr,
                            rm, grad, false);
}

hipError_t magnetization_forward(jgm_context* ctx, const double* model, double* out,
                                 double* raw_out) {
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long ncols = (long)((int)ctx->nx + 1) * ncy;
  const long nq = ncols * ncz;
  long nchunks = 0, per = 1;
  hipError_t e;
  if (N > 0) {
    site_chunks(S, ncols, &nchunks, &per);
    if ((e = ctx->weights.ensure(sizeof(double) * 3 * (size_t)nq)) != hipSuccess) return e;
    if ((e = ctx->partial.ensure(sizeof(double) * (size_t)nchunks * 3 * S)) != hipSuccess)
      return e;
    for (int j = 0; j < 3; ++j)
      hipLaunchKernelGGL(k_corner_weights, dim3(blocks_for(nq)), dim3(kBlock), 0, st,
                         model + j * N, 0L, 1, (int)ctx->nx, (int)ctx->ny, (int)ctx->nz,
                         ctx->weights.as<double>() + j * nq);
    ProfScope ps(ctx, "forward_corners_fp64_magnetization");
    hipLaunchKernelGGL(k_forward_corners_mag3, dim3((S + kBlock - 1) / kBlock, nchunks),
                       dim3(kBlock), 0, st, ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),
                       ctx->d_sz.as<double>(), S, ctx->d_Xc.as<double>(),
                       ctx->d_Yc.as<double>(), ctx->d_Zc.as<double>(), ncy, ncz,
                       ctx->weights.as<const double>(), nq, ncols, per,
                       ctx->partial.as<double>());
  }
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  if (nchunks > 0) {
    if ((e = ctx->partsum.ensure(sizeof(double) * 3 * (size_t)S)) != hipSuccess) return e;
    if ((e = launch_reduce_chunks(ctx, ctx->partial.as<const double>(), (int)nchunks, 3L * S,
                                  ctx->partsum.as<double>())) != hipSuccess)
      return e;
  }
  hipLaunchKernelGGL(k_mag3_finish, dim3(blocks_for(S)), dim3(kBlock), 0, st,
                     ctx->partsum.as<const double>(), nchunks > 0 ? 1 : 0, S, magnetization_scale(),
                     ctx->transform, tf_bx, tf_by, tf_bz, out, raw_out);
  return hipGetLastError();
}

hipError_t magnetization_adjoint(jgm_context* ctx, const double* rm, double* grad) {
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long nq = (long)ncorners(ctx);
  if (N == 0) return hipSuccess;
  hipError_t e;
  if ((e = ctx->coef.ensure(sizeof(double) * 3 * (size_t)S)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_mag3_coef, dim3(blocks_for(3L * S)), dim3(kBlock), 0, st, rm, S,
                     magnetization_scale(), ctx->coef.as<double>());
  int nchunks, per;
  corner_chunks(S, nq, &nchunks, &per);
  if ((e = ctx->hpart.ensure(sizeof(double) * (size_t)nchunks * 3 * nq)) != hipSuccess) return e;
  {
    ProfScope ps(ctx, "adjoint_corners_fp64_magnetization");
    hipLaunchKernelGGL(k_adjoint_corners_mag3, dim3(blocks_for(nq), nchunks), dim3(kBlock), 0, st,
                       ctx->d_sx.as<double>(), ctx->d_sy.as<double>(), ctx->d_sz.as<double>(),
                       ctx->coef.as<const double>(), S, per, ctx->d_Xc.as<double>(),
                       ctx->d_Yc.as<double>(), ctx->d_Zc.as<double>(), ncy, ncz, nq,
                       ctx->hpart.as<double>());
  }
  const double* H = ctx->hpart.as<const double>();
  if (nchunks > 1) {
    if ((e = ctx->hsum.ensure(sizeof(double) * 3 * (size_t)nq)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(3 * nq)), dim3(kBlock), 0, st, H, nchunks,
                       3 * nq, ctx->hsum.as<double>());
    H = ctx->hsum.as<const double>();
  }
  for (int j = 0; j < 3; ++j)
    hipLaunchKernelGGL(k_corner_difference, dim3(blocks_for(N)), dim3(kBlock), 0, st, H + j * nq,
                       (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, grad + j * N);
  return hipGetLastError();
}

}  // namespace

// out[i] = sum_c part[c * n + i].  A per-row loop over all chunks leaves the GPU idle
// when the rows are few (a forward over few sites has few rows and many chunks), so
// many chunks go through a grouped first pass (about 256k threads) and a second pass
// over the group totals.  The order of additions is fixed.
namespace {
// groups of the first pass for nchunks chunks of n rows (1: a single pass)
long reduce_groups(int nchunks, long n, int* per) {
  if (nchunks <= 8) {
    *per = nchunks;
    return 1;
  }
  long groups = std::max<long>(1, std::min<long>((262144 + n - 1) / n, nchunks / 4));
  *per = (int)((nchunks + groups - 1) / groups);
  return (nchunks + *per - 1) / *per;
}
// the reduction with the group totals (groups * n doubles) in `gsum`
void reduce_chunks_with(hipStream_t st, const double* part, int nchunks, long n, double* out,
                        double* gsum) {
  int per = 0;
  const long groups = reduce_groups(nchunks, n, &per);
  if (groups == 1) {
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(n)), dim3(kBlock), 0, st, part, nchunks, n,
                       out);
    return;
  }
  hipLaunchKernelGGL(k_sum_chunk_groups, dim3(blocks_for(n), (unsigned)groups), dim3(kBlock), 0,
                     st, part, nchunks, per, n, gsum);
  hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(n)), dim3(kBlock), 0, st,
                     static_cast<const double*>(gsum), (int)groups, n, out);
}
}  // namespace

hipError_t launch_reduce_chunks(jgm_context* ctx, const double* part, int nchunks, long n,
                                double* out) {
  if (n <= 0) return hipSuccess;
  int per = 0;
  const long groups = reduce_groups(nchunks, n, &per);
  if (groups > 1) {
    hipError_t e = ctx->chunk_scratch.ensure(sizeof(double) * (size_t)groups * n);
    if (e != hipSuccess) return e;
  }
  reduce_chunks_with(ctx->stream, part, nchunks, n, out, ctx->chunk_scratch.as<double>());
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 closed form for a subset of the sites: F sites by their coordinates (fx, fy, fz)
// and their indices idx in the survey.  The mixed-precision path hands these sites over
// (sites on mesh lines, where the reference's guarded logarithms differ from the field the
// far-field expansion computes); results go to the sites' own places in out / raw_out and
// their cell gradients are added to grad.
namespace {
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_forward_finish_sites(
    const double* __restrict__ psum, int F, const int* __restrict__ idx,
    const double* __restrict__ bgT, const double* __restrict__ bgvals, int L, int kind,
    FieldDir fd, int tf, double tfbx, double tfby, double tfbz, double* __restrict__ out,
    double* __restrict__ raw_out) {
  constexpr int NC = Ncomp<MASK>::value;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  double acc[NC], sc[NC];
  comp_scales<MASK>(sc);
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = psum[(long)m * F + f] * sc[m];
  double u[7];
  spread_components<MASK>(acc, u);
  const int s = idx[f];
  double raw[10];
  const int kraw = assemble_raw(u, s, bgT, bgvals, L, kind, fd, raw);
  store_site(raw, kraw, s, tf, tfbx, tfby, tfbz, out, raw_out);
}

// dst[f k + m] = src[idx[f] k + m]
__global__ __launch_bounds__(kBlock) void k_gather_site_rows(const double* __restrict__ src,
                                                             const int* __restrict__ idx, int F,
                                                             int k, double* __restrict__ dst) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)F * k) return;
  const int f = (int)(t / k), m = (int)(t % k);
  dst[t] = src[(long)idx[f] * k + m];
}

// rows idx[f] of a k-wide site array set to zero
__global__ __launch_bounds__(kBlock) void k_zero_site_rows(double* __restrict__ a,
                                                           const int* __restrict__ idx, int F,
                                                           int k) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)F * k) return;
  a[(long)idx[t / k] * k + t % k] = 0.0;
}

template <int MASK>
hipError_t forward_sites_impl(jgm_context* ctx, int kind, const double* model, int F,
                              const double* fx, const double* fy, const double* fz,
                              const int* idx, double* out, double* raw_out) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const int ncx = (int)ctx->nx + 1, ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long ncols = (long)ncx * ncy;
  const int site_blocks = (F + kBlock - 1) / kBlock;
  const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
  long want = std::min<long>((kTargetWaves + waves_per_chunk - 1) / waves_per_chunk, ncols);
  want = std::max<long>(1, want);
  const long per = (ncols + want - 1) / want;
  const int nchunks = (int)((ncols + per - 1) / per);
  hipError_t e;
  if ((e = ctx->rt_w.ensure(sizeof(double) * (size_t)ncols * ncz)) != hipSuccess ||
      (e = ctx->rt_part.ensure(sizeof(double) * (size_t)nchunks * NC * F)) != hipSuccess ||
      (e = ctx->rt_sum.ensure(sizeof(double) * (size_t)NC * F)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_corner_weights, dim3(blocks_for(ncols * ncz)), dim3(kBlock), 0, st, model,
                     0L, 1, (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, ctx->rt_w.as<double>());
  {
    ProfScope ps(ctx, "forward_corners_fp64_line_sites");
    launch_forward_corners<MASK>(1, dim3(site_blocks, nchunks), st, fx, fy, fz, F,
                                 ctx->d_Xc.as<const double>(), ctx->d_Yc.as<const double>(),
                                 ctx->d_Zc.as<const double>(), ncy, ncz,
                                 ctx->rt_w.as<const double>(), ncols, per,
                                 ctx->rt_part.as<double>());
  }
  if ((e = launch_reduce_chunks(ctx, ctx->rt_part.as<const double>(), nchunks, (long)NC * F,
                                ctx->rt_sum.as<double>())) != hipSuccess)
    return e;
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  hipLaunchKernelGGL((k_forward_finish_sites<MASK>), dim3(blocks_for(F)), dim3(kBlock), 0, st,
                     ctx->rt_sum.as<const double>(), F, idx, ctx->bgT.as<const double>(),
                     model + N, L, kind, field_dir(ctx), ctx->transform, tf_bx, tf_by, tf_bz,
                     out, raw_out);
  return hipGetLastError();
}

template <int MASK>
hipError_t adjoint_sites_impl(jgm_context* ctx, int kind, int F, const double* fx,
                              const double* fy, const double* fz, const int* idx,
                              const double* rm, double* grad) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const long N = (long)ncells(ctx);
  const int kraw = (int)raw_per_site(kind);
  const int ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long nq = (long)ncorners(ctx);
  const long corner_waves = (nq + 63) / 64;
  int nchunks = (int)std::min<long>(F, std::max<long>(1, (kTargetWaves + corner_waves - 1) /
                                                            corner_waves));
  const int per = (F + nchunks - 1) / nchunks;
  nchunks = (F + per - 1) / per;
  hipError_t e;
  if ((e = ctx->rt_rm.ensure(sizeof(double) * (size_t)F * kraw)) != hipSuccess ||
      (e = ctx->rt_coef.ensure(sizeof(double) * (size_t)F * NC)) != hipSuccess ||
      (e = ctx->rt_h.ensure(sizeof(double) * (size_t)nchunks * nq)) != hipSuccess ||
      (e = ctx->rt_sum.ensure(sizeof(double) * (size_t)std::max<long>(nq, N))) != hipSuccess ||
      (e = ctx->rt_tmp.ensure(sizeof(double) * (size_t)N)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_gather_site_rows, dim3(blocks_for((long)F * kraw)), dim3(kBlock), 0, st, rm,
                     idx, F, kraw, ctx->rt_rm.as<double>());
  hipLaunchKernelGGL((k_adjoint_coef<MASK>), dim3(blocks_for(F)), dim3(kBlock), 0, st,
                     ctx->rt_rm.as<const double>(), F, kind, field_dir(ctx),
                     ctx->rt_coef.as<double>(), NC);
  {
    ProfScope ps(ctx, "adjoint_corners_fp64_line_sites");
    launch_adjoint_corners<MASK>(1, dim3(blocks_for(nq), nchunks), st, fx, fy, fz,
                                 ctx->rt_coef.as<const double>(), F, per,
                                 ctx->d_Xc.as<const double>(), ctx->d_Yc.as<const double>(),
                                 ctx->d_Zc.as<const double>(), ncy, ncz, nq,
                                 ctx->rt_h.as<double>());
  }
  const double* H = ctx->rt_h.as<const double>();
  if (nchunks > 1) {
    hipLaunchKernelGGL(k_sum_chunks, dim3(blocks_for(nq)), dim3(kBlock), 0, st, H, nchunks, nq,
                       ctx->rt_sum.as<double>());
    H = ctx->rt_sum.as<const double>();
  }
  hipLaunchKernelGGL(k_corner_difference, dim3(blocks_for(N)), dim3(kBlock), 0, st, H,
                     (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, ctx->rt_tmp.as<double>());
  hipLaunchKernelGGL(k_add_f64, dim3(blocks_for(N)), dim3(kBlock), 0, st, grad,
                     ctx->rt_tmp.as<const double>(), N);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_forward_fp64_sites(jgm_context* ctx, int kind, const double* model, int F,
                                     const double* fx, const double* fy, const double* fz,
                                     const int* idx, double* out, double* raw_out) {
  if (F <= 0 || ncells(ctx) == 0) return hipSuccess;
  switch (mask_of(kind)) {
    case kMaskScalar:
      return forward_sites_impl<kMaskScalar>(ctx, kind, model, F, fx, fy, fz, idx, out, raw_out);
    case kMaskAll:
      return forward_sites_impl<kMaskAll>(ctx, kind, model, F, fx, fy, fz, idx, out, raw_out);
    default:
      return forward_sites_impl<kMaskTensor>(ctx, kind, model, F, fx, fy, fz, idx, out, raw_out);
  }
}

hipError_t launch_adjoint_fp64_sites_add(jgm_context* ctx, int kind, int F, const double* fx,
                                         const double* fy, const double* fz, const int* idx,
                                         const double* rm, double* grad) {
  if (F <= 0 || ncells(ctx) == 0) return hipSuccess;
  switch (mask_of(kind)) {
    case kMaskScalar:
      return adjoint_sites_impl<kMaskScalar>(ctx, kind, F, fx, fy, fz, idx, rm, grad);
    case kMaskAll: return adjoint_sites_impl<kMaskAll>(ctx, kind, F, fx, fy, fz, idx, rm, grad);
    default: return adjoint_sites_impl<kMaskTensor>(ctx, kind, F, fx, fy, fz, idx, rm, grad);
  }
}

hipError_t launch_maggrad_fp64_sites(jgm_context* ctx, const double* model, int F,
                                     const double* fx, const double* fy, const double* fz,
                                     const int* idx, double* out, double* raw_out) {
  if (F <= 0 || ncells(ctx) == 0) return hipSuccess;
  return maggrad_forward_at(ctx, model, F, fx, fy, fz, idx, out, raw_out);
}

hipError_t launch_maggrad_fp64_sites_add(jgm_context* ctx, int F, const double* fx,
                                         const double* fy, const double* fz, const int* idx,
                                         const double* rm, double* grad) {
  if (F <= 0 || ncells(ctx) == 0) return hipSuccess;
  return maggrad_adjoint_at(ctx, F, fx, fy, fz, idx, rm, grad, true);
}

hipError_t launch_zero_site_rows(jgm_context* ctx, double* a, const int* idx, int F, int k) {
  if (F <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_site_rows, dim3(blocks_for((long)F * k)), dim3(kBlock), 0,
                     ctx->stream, a, idx, F, k);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free FP64 forward (jgm_forward_noalloc).  Its working memory -- layer
// thicknesses, background table, corner weights, chunk partials and their sums -- lives
// in `scratch`, one file-scope array per GPU of JGM_NOALLOC_SCRATCH_BYTES, held until the
// process ends.  In host code `scratch` is a stand-in symbol, not a device address: the
// kernels get the address hipGetSymbolAddress returned for the device, looked up once per
// device when the first context on it is created.  Every region a kernel reads is written
// earlier in the same call (weights, table and partials cover their whole extent), so
// nothing needs clearing between calls; calls that overlap on one GPU race on it.
namespace {
constexpr size_t kNoallocDoubles = JGM_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
bool g_scratch_looked_up[kMaxScratchDevices] = {};
double* g_scratch_ptr[kMaxScratchDevices] = {};
hipError_t g_scratch_err[kMaxScratchDevices] = {};

// offsets in doubles rounded up to 256 bytes
inline size_t align32(size_t n) { return (n + 31) / 32 * 32; }

// Scratch layout of one no-alloc forward: [thick L][bgT 7 S L][weights ncorners]
// [partials nchunks NC S][partsum NC S][group sums]; 0 when even one chunk does not fit.
struct NoallocPlan {
  size_t thick = 0, bg = 0, w = 0, part = 0, psum = 0, gsum = 0, end = 0;
  int nchunks = 0;
  long cols_per_chunk = 0;
};

size_t noalloc_layout(NoallocPlan* p, size_t L, size_t S, size_t nq, size_t NC, int nchunks) {
  p->thick = 0;
  p->bg = align32(p->thick + L);
  p->w = align32(p->bg + 7 * S * L);
  p->part = align32(p->w + nq);
  p->psum = align32(p->part + (size_t)nchunks * NC * S);
  p->gsum = align32(p->psum + NC * S);
  int per = 0;
  const long groups = nchunks > 0 ? reduce_groups(nchunks, (long)(NC * S), &per) : 1;
  p->end = p->gsum + (groups > 1 ? (size_t)groups * NC * S : 0);
  p->nchunks = nchunks;
  return p->end;
}
}  // namespace

void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_looked_up[device]) return;
  void* ptr = nullptr;
  // the caller has made `device` current
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? static_cast<double*>(ptr) : nullptr;
  g_scratch_looked_up[device] = true;
}

double* noalloc_scratch(int device, hipError_t* err) {
  if (device < 0 || device >= kMaxScratchDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  *err = g_scratch_looked_up[device] ? g_scratch_err[device] : hipErrorNotInitialized;
  return g_scratch_ptr[device];
}

size_t noalloc_forward_need(const jgm_context* ctx, int kind) {
  NoallocPlan p;
  const size_t NC = mask_of(kind) == kMaskScalar ? Ncomp<kMaskScalar>::value
                    : mask_of(kind) == kMaskAll  ? Ncomp<kMaskAll>::value
                                                 : Ncomp<kMaskTensor>::value;
  return sizeof(double) * noalloc_layout(&p, ctx->bgthick.size(), ctx->S,
                                         ncells(ctx) > 0 ? ncorners(ctx) : 0, NC,
                                         ncells(ctx) > 0 ? 1 : 0);
}

namespace {
template <int MASK>
hipError_t forward_noalloc_impl(jgm_context* ctx, int kind, double* base, const double* model,
                                double* out) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const int ncx = (int)ctx->nx + 1, ncy = (int)ctx->ny + 1, ncz = (int)ctx->nz + 1;
  const long ncols = (long)ncx * ncy;
  const size_t nq = N > 0 ? (size_t)ncols * ncz : 0;
  const int site_blocks = (S + kBlock - 1) / kBlock;
  // as many chunks as forward_impl takes, or as many as fit
  NoallocPlan plan;
  int nchunks = 0;
  if (N > 0) {
    const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
    long want = std::min<long>((kTargetWaves + waves_per_chunk - 1) / waves_per_chunk, ncols);
    want = std::max<long>(1, want);
    while (want > 1 && noalloc_layout(&plan, L, S, nq, NC, (int)want) > kNoallocDoubles)
      want = want * 3 / 4;
    const long per = (ncols + want - 1) / want;
    nchunks = (int)((ncols + per - 1) / per);
    plan.cols_per_chunk = per;
  }
  if (noalloc_layout(&plan, L, S, nq, NC, nchunks) > kNoallocDoubles) return hipErrorOutOfMemory;
  double* thick = base + plan.thick;
  double* bgT = base + plan.bg;
  if (L > 0) {
    hipError_t e = hipMemcpyAsync(thick, ctx->bgthick.data(), sizeof(double) * L,
                                  hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_background_table, dim3(blocks_for(S)), dim3(kBlock), 0, st,
                       ctx->d_sx.as<double>(), ctx->d_sy.as<double>(), ctx->d_sz.as<double>(), S,
                       static_cast<const double*>(thick), L, ctx->xc[0], ctx->yc[0], ctx->zc[0],
                       ctx->xwidth, ctx->ywidth, ctx->zwidth, bgT);
  }
  if (nchunks > 0) {
    double* w = base + plan.w;
    double* part = base + plan.part;
    hipLaunchKernelGGL(k_corner_weights, dim3(blocks_for(ncols * ncz)), dim3(kBlock), 0, st,
                       model, 0L, 1, (int)ctx->nx, (int)ctx->ny, (int)ctx->nz, w);
    launch_forward_corners<MASK>(1, dim3(site_blocks, nchunks), st, ctx->d_sx.as<const double>(),
                                 ctx->d_sy.as<const double>(), ctx->d_sz.as<const double>(), S,
                                 ctx->d_Xc.as<const double>(), ctx->d_Yc.as<const double>(),
                                 ctx->d_Zc.as<const double>(), ncy, ncz,
                                 static_cast<const double*>(w), ncols, plan.cols_per_chunk, part);
    reduce_chunks_with(st, part, nchunks, (long)NC * S, base + plan.psum, base + plan.gsum);
  }
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  hipLaunchKernelGGL((k_forward_finish<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st,
                     static_cast<const double*>(base + plan.psum), nchunks > 0 ? 1 : 0, S,
                     static_cast<const double*>(bgT), model + N, L, kind, field_dir(ctx),
                     ctx->transform, tf_bx, tf_by, tf_bz, out, nullptr);
  return hipGetLastError();
}
}  // namespace

hipError_t launch_forward_fp64_noalloc(jgm_context* ctx, int kind, const double* model,
                                       double* out) {
  hipError_t e = hipSuccess;
  double* base = noalloc_scratch(ctx->device, &e);
  if (!base) return e == hipSuccess ? hipErrorInvalidValue : e;
  switch (mask_of(kind)) {
    case kMaskScalar: return forward_noalloc_impl<kMaskScalar>(ctx, kind, base, model, out);
    case kMaskAll: return forward_noalloc_impl<kMaskAll>(ctx, kind, base, model, out);
    default: return forward_noalloc_impl<kMaskTensor>(ctx, kind, base, model, out);
  }
}

namespace {
__global__ __launch_bounds__(kBlock) void k_zero_flag(int* __restrict__ flag) {
  if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 0;
}
// every lane that finds a differing element stores the same value 1
__global__ __launch_bounds__(kBlock) void k_models_differ(const double* __restrict__ a,
                                                          const double* __restrict__ b, long n,
                                                          int* __restrict__ flag) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    if (__double_as_longlong(a[i]) != __double_as_longlong(b[i])) flag[0] = 1;
}
}  // namespace

hipError_t launch_models_differ(jgm_context* ctx, const double* a, const double* b, size_t n,
                                int* flag) {
  hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(kBlock), 0, ctx->stream, flag);
  if (n > 0)
    hipLaunchKernelGGL(k_models_differ, dim3((unsigned)std::min<size_t>(blocks_for((long)n), 1024)),
                       dim3(kBlock), 0, ctx->stream, a, b, (long)n, flag);
  return hipGetLastError();
}

hipError_t launch_background_gradient(jgm_context* ctx, int kind, const double* rm,
                                      double* grad_bg) {
  const int L = (int)ctx->bgthick.size();
  if (L == 0) return hipSuccess;
  hipLaunchKernelGGL(k_background_gradient, dim3(L), dim3(kBlock), 0, ctx->stream, rm,
                     ctx->bgT.as<const double>(), (int)ctx->S, L, kind, field_dir(ctx), grad_bg);
  return hipGetLastError();
}

hipError_t launch_background_table(jgm_context* ctx) {
  const int S = (int)ctx->S;
  const int L = (int)ctx->bgthick.size();
  hipError_t e = ctx->bgT.ensure(sizeof(double) * (size_t)std::max(1, S) * std::max(1, L) * 7);
  if (e != hipSuccess) return e;
  if (L == 0 || S == 0) {
    ctx->bg_valid = true;
    return hipSuccess;
  }
  DeviceBuffer thick;
  e = thick.ensure(sizeof(double) * L);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(thick.ptr, ctx->bgthick.data(), sizeof(double) * L, hipMemcpyHostToDevice,
                     ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_background_table, dim3(blocks_for(S)), dim3(kBlock), 0, ctx->stream,
                       ctx->d_sx.as<double>(), ctx->d_sy.as<double>(), ctx->d_sz.as<double>(), S,
                       thick.as<const double>(), L, ctx->xc[0], ctx->yc[0], ctx->zc[0],
                       ctx->xwidth, ctx->ywidth, ctx->zwidth, ctx->bgT.as<double>());
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  thick.release();
  if (e != hipSuccess) return e;
  if (e2 != hipSuccess) return e2;
  ctx->bg_valid = true;
  return hipSuccess;
}

hipError_t launch_forward_fp64(jgm_context* ctx, int kind, const double* model, double* out,
                               double* raw_out) {
  if (kind == JGM_KIND_MAGNETIC_GRADIENT) return maggrad_forward(ctx, model, out, raw_out);
  if (kind == JGM_KIND_MAGNETIZATION) return magnetization_forward(ctx, model, out, raw_out);
  return launch_forward_fp64_batch(ctx, kind, 1, model, 0, out, 0, raw_out, 0);
}

hipError_t launch_forward_fp64_batch(jgm_context* ctx, int kind, size_t nmodels,
                                     const double* models, size_t model_ld, double* out,
                                     size_t out_ld, double* raw_out, size_t raw_ld) {
  for (size_t b0 = 0; b0 < nmodels;) {
    const size_t left = nmodels - b0;
    const int nb = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const double* m = models + b0 * model_ld;
    double* o = out + b0 * out_ld;
    double* r = raw_out ? raw_out + b0 * raw_ld : nullptr;
    hipError_t e;
    if (kind == JGM_KIND_MAGNETIC_GRADIENT || kind == JGM_KIND_MAGNETIZATION) {
      e = kind == JGM_KIND_MAGNETIC_GRADIENT ? maggrad_forward(ctx, m, o, r)
                                             : magnetization_forward(ctx, m, o, r);
      b0 += 1;
    } else {
      switch (mask_of(kind)) {
        case kMaskScalar:
          e = forward_impl<kMaskScalar>(ctx, kind, nb, m, model_ld, o, out_ld, r, raw_ld);
          break;
        case kMaskAll:
          e = forward_impl<kMaskAll>(ctx, kind, nb, m, model_ld, o, out_ld, r, raw_ld);
          break;
        default:
          e = forward_impl<kMaskTensor>(ctx, kind, nb, m, model_ld, o, out_ld, r, raw_ld);
          break;
      }
      b0 += nb;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_misfit_to_raw(jgm_context* ctx, int kind, const double* misfit,
                                const double* raw, double* raw_misfit) {
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  hipLaunchKernelGGL(k_misfit_to_raw, dim3(blocks_for((long)ctx->S)), dim3(kBlock), 0,
                     ctx->stream, misfit, raw, (int)ctx->S, (int)raw_per_site(kind),
                     ctx->transform, tf_bx, tf_by, tf_bz, raw_misfit);
  return hipGetLastError();
}

hipError_t launch_adjoint_fp64(jgm_context* ctx, int kind, const double* /*model*/,
                               const double* raw_misfit, double* grad) {
  return launch_adjoint_fp64_batch(ctx, kind, 1, raw_misfit, 0, grad, 0);
}

hipError_t launch_adjoint_fp64_batch(jgm_context* ctx, int kind, size_t nsets,
                                     const double* raw_misfits, size_t misfit_ld, double* grads,
                                     size_t grad_ld) {
  for (size_t b0 = 0; b0 < nsets;) {
    const size_t left = nsets - b0;
    const int nb = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const double* r = raw_misfits + b0 * misfit_ld;
    double* g = grads + b0 * grad_ld;
    hipError_t e;
    if (kind == JGM_KIND_MAGNETIC_GRADIENT || kind == JGM_KIND_MAGNETIZATION) {
      e = kind == JGM_KIND_MAGNETIC_GRADIENT ? maggrad_adjoint(ctx, r, g)
                                             : magnetization_adjoint(ctx, r, g);
      b0 += 1;
    } else {
      switch (mask_of(kind)) {
        case kMaskScalar:
          e = adjoint_impl<kMaskScalar>(ctx, kind, nb, r, misfit_ld, g, grad_ld);
          break;
        case kMaskAll:
          e = adjoint_impl<kMaskAll>(ctx, kind, nb, r, misfit_ld, g, grad_ld);
          break;
        default:
          e = adjoint_impl<kMaskTensor>(ctx, kind, nb, r, misfit_ld, g, grad_ld);
          break;
      }
      b0 += nb;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_sensitivity_rows(jgm_context* ctx, int kind, size_t site0, size_t nsites,
                                   double* rows) {
  const long np = (long)nparams(ctx, kind);
  const int L = (int)ctx->bgthick.size();
  // unit factors of the vertical-gradient and magnetization rows (the terms include G)
  const double mscale = kind == JGM_KIND_MAGNETIZATION
                            ? magnetization_scale() / kGravConst
                            : ctx->fieldstrength * (1e-7 / kGravConst);
  dim3 grid(blocks_for(np), (unsigned)nsites);
  const double* a[6] = {ctx->d_xlo.as<double>(), ctx->d_ylo.as<double>(), ctx->d_zlo.as<double>(),
                        ctx->d_xsz.as<double>(), ctx->d_ysz.as<double>(), ctx->d_zsz.as<double>()};
#define JGM_ROWS(M)                                                                              \
  hipLaunchKernelGGL((k_sensitivity_rows<M>), grid, dim3(kBlock), 0, ctx->stream,                \
                     ctx->d_sx.as<double>(), ctx->d_sy.as<double>(), ctx->d_sz.as<double>(),     \
                     (int)site0, a[0], a[1], a[2], a[3], a[4], a[5], (int)ctx->nx, (int)ctx->ny, \
                     (int)ctx->nz, ctx->bgT.as<const double>(), L, kind, field_dir(ctx), mscale, \
                     rows)
  switch (mask_of(kind)) {
    case kMaskScalar: JGM_ROWS(kMaskScalar); break;
    case kMaskAll: JGM_ROWS(kMaskAll); break;
    default: JGM_ROWS(kMaskTensor); break;
  }
#undef JGM_ROWS
  return hipGetLastError();
}

}  // namespace jgm
