// Mixed-precision path.
//
// Split per site: cells inside the site's near box (every cell closer than q0
// cell half-diagonals lies inside it) use the FP64 closed form; all other cells use
// the FP32 multipole expansion of the uniform prism about its centre through the
// hexadecapole:
//
//   Phi/V = 1/r + Q(x)/r^5 + H(x)/r^9,
//   Q = sum_m B_m x_m^2,           B_m = (h_m^2 - hbar^2)/2, hbar^2 = mean of h^2,
//   H = harmonic part of <(u.x)^4> over the box times 35/8
//     = a1 x^4 + a2 y^4 + a3 z^4 + b12 x^2 y^2 + b13 x^2 z^2 + b23 y^2 z^2,
//
// with x = site - cell centre and h the half sizes; gz = G rho V dPhi/dz and
// U_ij = G rho V d2Phi/dx_i dx_j.  Odd orders vanish by symmetry, so the neglected
// term is O((h/d)^6) relative.  Cubes have B = 0.  The formulas are evaluated in
// direction cosines u = x/r so that no power of 1/r below r^-7 is formed (FP32
// range); 1/r is the hardware estimate, refined by one Newton step in the multipole
// path and used as it is, with its mean bias taken out in FP64, in the monopole paths.
// Beyond a second radius (q1 half-diagonals of the tile's largest cell, default 28,
// tested per tile against the distance from the site to the tile's nearest cell centre) the
// hexadecapole, O((h/d)^4) of a cell's field there, is dropped; for cubes the remaining
// monopole is evaluated in a
// regrouped form (running sums of w s^3 and w s^5 along a column of cells in the
// forward, the site's coefficients folded into one quadratic form in the adjoint).
// Site offsets are formed in FP64 against the first cell of a tile and rounded once
// per tile (forward) or once per site and lane (adjoint).  Accumulation per site
// (forward) and per cell (adjoint) runs in FP32 over short runs, flushed to FP64 every
// 4 cells or 4 columns (forward) or 8 sites (adjoint): the rounding error of an FP32
// run grows with the run length times its partial sum.  Forward waves take their 64
// sites in Morton order of (x, y) so that a wave's sites cluster and tiles outside all
// their near boxes and second radii take the cheap path for the whole wave.
//
// Layouts in HBM
//   cell table   FP32: uniform mesh [c] = {w}; otherwise [c][12] =
//                {w, Bx, By, Bz, 4 a1, 4 a2, 4 a3, 2 b12, 2 b13, 2 b23, 0, 0}
//                w = G rho V (forward) or G V (adjoint)
//   far partials forward [chunk][NC][S] FP64; adjoint [chunk][N] FP64
//   near boxes   int32 [S][6]; per-column CSR of sites; per-site offsets of the
//                near corner values and near cell values of the adjoint
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

#include "context.hpp"
#include "epilogue.hpp"
#include "prism_device.hpp"

namespace jgm {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 4;  // far-field tile edge (cells)
// Waves per far forward launch (jgm_context::far_fwd_waves, 65536 by default): the tiles are
// cut into equal chunks to reach about this many.  The tiles' costs differ (monopole, hexadecapole, masked near-box tiles), and many
// small chunks balance the CUs: at 128x128x64 cells and 4096 sites, measured against 16384,
// the forward is fastest at 65536 (2.40 -> 2.04 ms); chunks of equal estimated cost per
// block of 256 sites (as the adjoint's partition below, from the paths each wave takes per
// tile) were slower at every count tried (8192-65536 waves: 1.73-2.24 ms against 1.69 ms for
// these equal chunks, which hold whole columns of tiles).  The partial sums this leaves
// (~0.3 GB per model) are reduced in fixed order.
constexpr int kTableWidth = 12;  // floats per cell of a non-uniform mesh

// Mean relative error of the hardware reciprocal square root (v_rsq_f32) on gfx950,
// measured over 4M values of r^2 log-uniform in [1e2, 1e7]: -7.4e-9 (-8.6e-9 for even
// binary exponents of r^2, -6.3e-9 for odd), rms 2.8e-8 -- about the rms that one
// Newton step leaves, but biased.  The monopole paths (cells beyond the second radius,
// most pairs) use the estimate without a Newton step and take the bias out of their
// FP32 sums in FP64: s^3 and s^5 are low by 3 and 5 times it, coherently over millions
// of cells.  The multipole path, whose terms are larger, keeps the Newton step.
constexpr double kRsqBias = 7.416e-9;
constexpr double kFixS3 = 1.0 + 3.0 * kRsqBias;
constexpr double kFixS5 = 1.0 + 5.0 * kRsqBias;

// Mesh classes of the far-field kernels.
enum : int { kMeshCubes = 0, kMeshUniform = 1, kMeshGeneral = 2 };

inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__global__ void k_fill_zero(double* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Multipole constants of one prism shape (see the file comment), stored as the products the
// field formulas use: B_m, 4 a_m and 2 b_mn.
struct MpConst {
  float Bx, By, Bz, a4x, a4y, a4z, b2xy, b2xz, b2yz;
};

__host__ __device__ inline MpConst mp_const(double dx, double dy, double dz) {
  const double hx2 = 0.25 * dx * dx, hy2 = 0.25 * dy * dy, hz2 = 0.25 * dz * dz;
  const double hb = (hx2 + hy2 + hz2) / 3.0;
  // fourth moments <u^4> = h^4/5 and multinomial x^2 y^2 terms 6 <u^2><v^2> = 2 h^2 k^2 / 3
  const double A = hx2 * hx2 / 5, B = hy2 * hy2 / 5, C = hz2 * hz2 / 5;
  const double D = 2 * hx2 * hy2 / 3, E = 2 * hx2 * hz2 / 3, F = 2 * hy2 * hz2 / 3;
  // harmonic projection P - r^2 lap(P)/14 + r^4 lap^2(P)/280
  const double p = 12 * A + 2 * D + 2 * E, q = 12 * B + 2 * D + 2 * F, t = 12 * C + 2 * E + 2 * F;
  const double w = 2 * (p + q + t);
  const double c = 35.0 / 8.0;  // (2n-1)!!/n! for n = 4
  MpConst m;
  m.Bx = (float)(0.5 * (hx2 - hb));
  m.By = (float)(0.5 * (hy2 - hb));
  m.Bz = (float)(0.5 * (hz2 - hb));
  m.a4x = (float)(4 * c * (A - p / 14 + w / 280));
  m.a4y = (float)(4 * c * (B - q / 14 + w / 280));
  m.a4z = (float)(4 * c * (C - t / 14 + w / 280));
  m.b2xy = (float)(2 * c * (D - (p + q) / 14 + w / 140));
  m.b2xz = (float)(2 * c * (E - (p + t) / 14 + w / 140));
  m.b2yz = (float)(2 * c * (F - (q + t) / 14 + w / 140));
  return m;
}

// Field of a prism of unit G*rho*V at offset (X, Y, Z) = site - centre, in the
// kernel component order of MASK.  QUAD = false drops the quadrupole (B = 0 for
// cubes); HEXA = false drops the hexadecapole (used beyond the second radius).
// With v the squared direction cosines, h_x = dH/dv_x = 4 a1 vx + 2 b12 vy + 2 b13 vz (and
// likewise h_y, h_z), and H is homogeneous of degree 2 in v, so 4 H = vx h_x + vy h_y + vz h_z:
// the hexadecapole value comes from its gradient in three operations.
template <int MASK, bool QUAD, bool HEXA = true>
__device__ __forceinline__ void multipole_f32(float X, float Y, float Z, const MpConst& c,
                                              float* T) {
  const float r2 = fmaf(X, X, fmaf(Y, Y, Z * Z));
  float s = __builtin_amdgcn_rsqf(r2);
  const float e = fmaf(-(r2 * s), s, 1.0f);  // one Newton step on the hardware estimate
  s = fmaf(0.5f * s, e, s);
  const float s2 = s * s;
  const float s3 = s2 * s;
  const float s5 = s3 * s2;
  const float s7 = s5 * s2;
  const float ux = X * s, uy = Y * s, uz = Z * s;
  const float vx = ux * ux, vy = uy * uy, vz = uz * uz;
  float P = 3.0f * s3;  // coefficient of u_i u_j
  float base = -s3;     // isotropic diagonal part
  float hx = 0.0f, hy = 0.0f, hz = 0.0f, tx = 0.0f, ty = 0.0f, tz = 0.0f;
  if constexpr (HEXA) {
    tx = c.a4x * vx;
    ty = c.a4y * vy;
    tz = c.a4z * vz;
    hx = fmaf(c.b2xy, vy, fmaf(c.b2xz, vz, tx));
    hy = fmaf(c.b2xy, vx, fmaf(c.b2yz, vz, ty));
    hz = fmaf(c.b2xz, vx, fmaf(c.b2yz, vy, tz));
    const float hh4 = fmaf(vx, hx, fmaf(vy, hy, vz * hz));  // 4 H
    P = fmaf(24.75f * hh4, s7, P);                            // + 99 H s^7
    base = fmaf(-2.25f * hh4, s7, base);                      // - 9 H s^7
  }
  if constexpr (QUAD) {
    const float q = fmaf(c.Bx, vx, fmaf(c.By, vy, c.Bz * vz));
    P = fmaf(35.0f * q, s5, P);
    base = fmaf(-5.0f * q, s5, base);
  }
  const float t2 = 2.0f * s5;
  int k = 0;
  if constexpr ((MASK & kGz) != 0) {
    float g = base;
    if constexpr (HEXA) g = fmaf(hz, s7, g);
    if constexpr (QUAD) g = fmaf(c.Bz, t2, g);
    T[k++] = Z * g;
  }
  if constexpr ((MASK & kDiag) != 0) {
    float px = P, py = P, pz = P, ox = base, oy = base, oz = base;
    if constexpr (HEXA) {
      // h_xx = h_x + 8 a1 vx = h_x + 2 (4 a1 vx), likewise for y and z
      const float m18 = -18.0f * s7;
      px = fmaf(hx, m18, px); py = fmaf(hy, m18, py); pz = fmaf(hz, m18, pz);
      ox = fmaf(fmaf(2.0f, tx, hx), s7, ox);
      oy = fmaf(fmaf(2.0f, ty, hy), s7, oy);
      oz = fmaf(fmaf(2.0f, tz, hz), s7, oz);
    }
    if constexpr (QUAD) {
      const float m20 = -20.0f * s5;
      px = fmaf(c.Bx, m20, px); py = fmaf(c.By, m20, py); pz = fmaf(c.Bz, m20, pz);
      ox = fmaf(c.Bx, t2, ox); oy = fmaf(c.By, t2, oy); oz = fmaf(c.Bz, t2, oz);
    }
    T[k++] = fmaf(vx, px, ox);
    T[k++] = fmaf(vy, py, oy);
    T[k++] = fmaf(vz, pz, oz);
  }
  if constexpr ((MASK & kOff) != 0) {
    float cxy = P, cxz = P, cyz = P;
    if constexpr (HEXA) {
      const float m9 = -9.0f * s7, f2 = 2.0f * s7;  // 4 b s^7 = (2 b) (2 s^7)
      cxy = fmaf(hx + hy, m9, fmaf(c.b2xy, f2, P));
      cxz = fmaf(hx + hz, m9, fmaf(c.b2xz, f2, P));
      cyz = fmaf(hy + hz, m9, fmaf(c.b2yz, f2, P));
    }
    if constexpr (QUAD) {
      const float m10 = -10.0f * s5;
      cxy = fmaf(c.Bx + c.By, m10, cxy);
      cxz = fmaf(c.Bx + c.Bz, m10, cxz);
      cyz = fmaf(c.By + c.Bz, m10, cyz);
    }
    T[k++] = (ux * uy) * cxy;
    T[k++] = (ux * uz) * cxz;
    T[k++] = (uy * uz) * cyz;
  }
  if constexpr ((MASK & kZcol) != 0) {
    // Uzz, Uxz, Uyz by the same operations as the two branches above
    float pz = P, oz = base, cxz = P, cyz = P;
    if constexpr (HEXA) {
      pz = fmaf(hz, -18.0f * s7, pz);
      oz = fmaf(fmaf(2.0f, tz, hz), s7, oz);
      const float m9 = -9.0f * s7, f2 = 2.0f * s7;
      cxz = fmaf(hx + hz, m9, fmaf(c.b2xz, f2, P));
      cyz = fmaf(hy + hz, m9, fmaf(c.b2yz, f2, P));
    }
    if constexpr (QUAD) {
      pz = fmaf(c.Bz, -20.0f * s5, pz);
      oz = fmaf(c.Bz, t2, oz);
      const float m10 = -10.0f * s5;
      cxz = fmaf(c.Bx + c.Bz, m10, cxz);
      cyz = fmaf(c.By + c.Bz, m10, cyz);
    }
    T[k++] = fmaf(vz, pz, oz);
    T[k++] = (ux * uz) * cxz;
    T[k++] = (uy * uz) * cyz;
  }
}

// multipole_f32 for a cube (no quadrupole) with the hexadecapole.  The cube's fourth-moment
// constants are a1 = a2 = a3 = a and b12 = b13 = b23 = -3a (mp_const; -3a makes H harmonic),
// so in direction cosines v (vx + vy + vz = 1) with sigma = vx vy + vx vz + vy vz:
//   H = a (1 - 5 sigma),  h_x = a (10 vx - 6),  h_xx = a (18 vx - 6),
//   off-diagonal factor P + 4 b s^7 - 9 (h_x + h_y) s^7 = P + a s^7 (6 + 90 vz),
// the same terms as the general form, with about a fifth fewer operations.
template <int MASK>
__device__ __forceinline__ void multipole_cube_f32(float X, float Y, float Z, float a, float* T) {
  const float r2 = fmaf(X, X, fmaf(Y, Y, Z * Z));
  float s = __builtin_amdgcn_rsqf(r2);
  const float e = fmaf(-(r2 * s), s, 1.0f);  // one Newton step on the hardware estimate
  s = fmaf(0.5f * s, e, s);
  const float s2 = s * s;
  const float s3 = s2 * s;
  const float s7 = (s3 * s2) * s2;
  const float ux = X * s, uy = Y * s, uz = Z * s;
  const float vx = ux * ux, vy = uy * uy, vz = uz * uz;
  const float as = a * s7;
  const float sig = fmaf(vx, vy, fmaf(vx, vz, vy * vz));
  const float K = fmaf(-5.0f * as, sig, as);  // H s^7
  const float P = fmaf(99.0f, K, 3.0f * s3);  // coefficient of u_i u_j
  const float base = fmaf(-9.0f, K, -s3);     // isotropic diagonal part
  const float u6 = 6.0f * as, u10 = 10.0f * as;
  int k = 0;
  if constexpr ((MASK & kGz) != 0) T[k++] = Z * (base + fmaf(u10, vz, -u6));
  if constexpr ((MASK & kDiag) != 0) {
    const float u18 = 18.0f * as, bm6 = base - u6;
    const float hxs = fmaf(u10, vx, -u6), hys = fmaf(u10, vy, -u6), hzs = fmaf(u10, vz, -u6);
    T[k++] = fmaf(vx, fmaf(-18.0f, hxs, P), fmaf(u18, vx, bm6));
    T[k++] = fmaf(vy, fmaf(-18.0f, hys, P), fmaf(u18, vy, bm6));
    T[k++] = fmaf(vz, fmaf(-18.0f, hzs, P), fmaf(u18, vz, bm6));
  }
  if constexpr ((MASK & kOff) != 0) {
    const float u90 = 90.0f * as, Pu6 = P + u6;
    T[k++] = (ux * uy) * fmaf(u90, vz, Pu6);
    T[k++] = (ux * uz) * fmaf(u90, vy, Pu6);
    T[k++] = (uy * uz) * fmaf(u90, vx, Pu6);
  }
}

template <int MESH>
__device__ __forceinline__ MpConst cell_const(const float* __restrict__ table, long c,
                                              const MpConst& uniform) {
  if constexpr (MESH != kMeshGeneral) {
    return uniform;
  } else {
    const float* t = table + c * kTableWidth + 1;
    MpConst m;
    m.Bx = t[0]; m.By = t[1]; m.Bz = t[2];
    m.a4x = t[3]; m.a4y = t[4]; m.a4z = t[5];
    m.b2xy = t[6]; m.b2xz = t[7]; m.b2yz = t[8];
    return m;
  }
}

// Cell table: w = G * rho * V (rho = 1 when rho is null); the shape constants
// for non-uniform meshes.
template <int MESH>
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
  const double w = kGravConst * (rho ? rho[c] : 1.0) * (dx * dy * dz);
  if constexpr (MESH != kMeshGeneral) {
    table[c] = (float)w;
  } else {
    const MpConst m = mp_const(dx, dy, dz);
    float* o = table + c * kTableWidth;
    o[0] = (float)w;
    o[1] = m.Bx; o[2] = m.By; o[3] = m.Bz;
    o[4] = m.a4x; o[5] = m.a4y; o[6] = m.a4z;
    o[7] = m.b2xy; o[8] = m.b2xz; o[9] = m.b2yz;
    o[10] = 0.0f; o[11] = 0.0f;
  }
}

// Cell weights of nb models: wt[c * nb + b] = G rho_b V (FP32), rho_b at rho + b * model_ld.
__global__ __launch_bounds__(kBlock) void k_cell_weights(const double* __restrict__ rho,
                                                         long model_ld, int nb,
                                                         const double* __restrict__ xsz,
                                                         const double* __restrict__ ysz,
                                                         const double* __restrict__ zsz, int nx,
                                                         int ny, int nz, float* __restrict__ wt) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n = (long)nx * ny * nz;
  if (c >= n) return;
  const int k = (int)(c % nz);
  const long t = c / nz;
  const int j = (int)(t % ny);
  const int i = (int)(t / ny);
  const double vol = xsz[i] * ysz[j] * zsz[k];
  for (int b = 0; b < nb; ++b) wt[c * nb + b] = (float)(kGravConst * rho[b * model_ld + c] * vol);
}

// The far field's geometry.  Along z it runs over nzf far layers: every mesh layer, or
// a listed subset (kmap[f] the mesh layer of far layer f, increasing); zcen, zoff, hz2 and
// hz2_16 are indexed by far layer, tiles group runs of far layers, and cell indices and
// near-box tests use mesh layers (nz of them).
struct FarGeom {
  const double* xcen;
  const double* ycen;
  const double* zcen;
  const float* xoff;  // centre minus the centre of the first cell of its 4-run (FP32)
  const float* yoff;
  const float* zoff;
  // squared largest half-size per 4-run of cells along each axis (z also per 16-run, the
  // adjoint's tile depth): a tile's squared largest half-diagonal is the sum of its three
  const float* hx2;
  const float* hy2;
  const float* hz2;
  const float* hz2_16;
  const int* kmap;  // null: far layer f is mesh layer f
  int nx, ny, nz, nzf;
  int ntx, nty, ntz;
};

// Mesh layer of far layer f (never listed on meshes of cubes, whose far field keeps every
// layer)
template <int MESH>
__device__ __forceinline__ int mesh_layer(const FarGeom& g, int f) {
  if constexpr (MESH == kMeshCubes)
    return f;
  else
    return g.kmap ? g.kmap[f] : f;
}

template <int MESH>
struct TableWidth {
  static constexpr int value = MESH == kMeshGeneral ? kTableWidth : 1;
};

// Sum over the cells of one 4x4x4 tile for one lane (site at Xt, Yt, Zt from the
// centre of the tile's first cell), flushed to FP64 per column of cells, for NB models
// at once (weights wt[c * NB + b]; the multipole terms of a pair are evaluated once).
// MASKED skips the lane's near-box cells (evaluated at a harmless far point, weight 0).
// shape: the unit-density cell table holding the shape constants of general meshes.
template <int MASK, int MESH, bool HEXA, bool MASKED, int NB>
__device__ __forceinline__ void far_tile(const FarGeom& g, const float* __restrict__ shape,
                                         const float* __restrict__ wt, const MpConst& uc,
                                         int i0, int j0, int k0, int ni, int nj, int nk,
                                         float Xt, float Yt, float Zt, const int (&bx)[6],
                                         double (&acc64)[NB][Ncomp<MASK>::value]) {
  constexpr int NC = Ncomp<MASK>::value;
  constexpr bool QUAD = MESH != kMeshCubes;
  for (int ii = 0; ii < ni; ++ii) {
    const float X = Xt - g.xoff[i0 + ii];
    const bool inx = MASKED && (i0 + ii >= bx[0]) && (i0 + ii <= bx[1]);
    for (int jj = 0; jj < nj; ++jj) {
      const float Y = Yt - g.yoff[j0 + jj];
      const bool inxy = inx && (j0 + jj >= bx[2]) && (j0 + jj <= bx[3]);
      const long cb = ((long)(i0 + ii) * g.ny + (j0 + jj)) * g.nz;  // the column's layer 0
      float acc[NB][NC];
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int m = 0; m < NC; ++m) acc[b][m] = 0.0f;
      for (int kk = 0; kk < nk; ++kk) {
        const int km = mesh_layer<MESH>(g, k0 + kk);
        const bool near = inxy && (km >= bx[4]) && (km <= bx[5]);
        const float Z = near ? 1.0e6f : Zt - g.zoff[k0 + kk];
        const MpConst cc = cell_const<MESH>(shape, cb + km, uc);
        float T[NC];
        if constexpr (MESH == kMeshCubes && HEXA)
          multipole_cube_f32<MASK>(X, Y, Z, 0.25f * cc.a4x, T);
        else
          multipole_f32<MASK, QUAD, HEXA>(X, Y, Z, cc, T);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const float w = near ? 0.0f : wt[(cb + km) * NB + b];
#pragma unroll
          for (int m = 0; m < NC; ++m) acc[b][m] = fmaf(w, T[m], acc[b][m]);
        }
      }
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int m = 0; m < NC; ++m) acc64[b][m] += (double)acc[b][m];
    }
  }
}

// Monopole-only sum over a tile of cubes.  With s = 1/r and w the cell weights, the
// monopole terms of multipole_f32 regrouped are
//   gz = -sum w Z s^3,  Txx = 3X^2 sum w s^5 - sum w s^3,  Tyy = 3 sum w Y^2 s^5 - sum w s^3,
//   Tzz = 3 sum w Z^2 s^5 - sum w s^3,  Txy = 3X sum w Y s^5,  Txz = 3X sum w Z s^5,
//   Tyz = 3 sum w Y Z s^5,
// so a row of the tile (fixed X) needs seven running sums: the Y-free ones accumulate per
// cell, the Y-weighted ones per column from its sums of w s^5 and w Z s^5; sum w s^3 is
// not summed but formed as X^2 sum w s^5 + sum w Y^2 s^5 + sum w Z^2 s^5 (r^2 s^5 = s^3),
// which makes the diagonal exactly trace-free, Txx = 2 X^2 A5 - AYY5 - AZZ5 and so on.
// The components are formed once per row, about half the operations per cell of the full
// multipole.  The NB models share s.
template <int MASK, int NB>
__device__ __forceinline__ void far_tile_mono(const FarGeom& g, const float* __restrict__ wt,
                                              int i0, int j0, int k0, int ni, int nj, int nk,
                                              float Xt, float Yt, float Zt,
                                              double (&acc64)[NB][Ncomp<MASK>::value]) {
  constexpr bool GZ = (MASK & kGz) != 0, DIAG = (MASK & kDiag) != 0, OFF = (MASK & kOff) != 0;
  // z offsets of the tile's cells; a cell past the mesh edge sits far away with weight 0
  float Zc[kTile];
  bool kv[kTile];
#pragma unroll
  for (int kk = 0; kk < kTile; ++kk) {
    kv[kk] = kk < nk;
    Zc[kk] = kv[kk] ? Zt - g.zoff[k0 + kk] : 1.0e6f;
  }
  for (int ii = 0; ii < ni; ++ii) {
    const float X = Xt - g.xoff[i0 + ii];
    // FP32 sums over this row's cells, flushed to FP64 per row
    float A5[NB], AY5[NB], AYY5[NB], AZ3[NB], AZZ5[NB], AZ5[NB], AYZ5[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b)
      A5[b] = AY5[b] = AYY5[b] = AZ3[b] = AZZ5[b] = AZ5[b] = AYZ5[b] = 0.0f;
    for (int jj = 0; jj < nj; ++jj) {
      const float Y = Yt - g.yoff[j0 + jj];
      const float YY = Y * Y;
      const float XXYY = fmaf(X, X, YY);
      const float* wc = wt + (((long)(i0 + ii) * g.ny + (j0 + jj)) * g.nz + k0) * NB;
      float S5[NB], SZ5[NB];  // this column's sums of w s^5 and w Z s^5
#pragma unroll
      for (int kk = 0; kk < kTile; ++kk) {
        const float Z = Zc[kk];
        const float ZZ = Z * Z;
        const float r2 = XXYY + ZZ;
        const float sv = __builtin_amdgcn_rsqf(r2);  // bias removed at the row's flush
        const float s2 = sv * sv;
        const float s3 = s2 * sv;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const float w = kv[kk] ? wc[kk * NB + b] : 0.0f;
          const float q3 = w * s3;
          if constexpr (GZ) AZ3[b] = fmaf(Z, q3, AZ3[b]);
          if constexpr (DIAG || OFF) {
            const float q5 = q3 * s2;
            // the column sums start at their first cell's terms (no additions of zero)
            S5[b] = kk == 0 ? q5 : S5[b] + q5;
            if constexpr (DIAG) AZZ5[b] = fmaf(ZZ, q5, AZZ5[b]);
            if constexpr (OFF) SZ5[b] = kk == 0 ? Z * q5 : fmaf(Z, q5, SZ5[b]);
          }
        }
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if constexpr (DIAG || OFF) A5[b] += S5[b];
        if constexpr (DIAG) AYY5[b] = fmaf(YY, S5[b], AYY5[b]);
        if constexpr (OFF) {
          AY5[b] = fmaf(Y, S5[b], AY5[b]);
          AZ5[b] += SZ5[b];
          AYZ5[b] = fmaf(Y, SZ5[b], AYZ5[b]);
        }
      }
    }
    const float X3 = 3.0f * X;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int k = 0;
      // explicit FMAs: the same rounding in every instantiation (batches equal singles)
      auto put = [&](double fix, float v) {
        acc64[b][k] = fma(fix, (double)v, acc64[b][k]);
        ++k;
      };
      if constexpr (GZ) put(kFixS3, -AZ3[b]);
      if constexpr (DIAG) {
        const float P = (X * X) * A5[b];
        put(kFixS5, fmaf(2.0f, P, -(AYY5[b] + AZZ5[b])));
        put(kFixS5, fmaf(2.0f, AYY5[b], -(P + AZZ5[b])));
        put(kFixS5, fmaf(2.0f, AZZ5[b], -(P + AYY5[b])));
      }
      if constexpr (OFF) {
        put(kFixS5, X3 * AY5[b]);
        put(kFixS5, X3 * AZ5[b]);
        put(kFixS5, 3.0f * AYZ5[b]);
      }
    }
  }
}

// Forward far field: one lane per site; each wave walks a chunk of 4x4x4 tiles.
// Lanes take sites in `order` (spatially compact waves).  Cells inside the lane's
// near box contribute nothing here (FP64 near pass).  A tile whose cells are all beyond
// the second radius of every site of the wave keeps only the monopole (and quadrupole)
// terms.  Site offsets are formed in FP64 against the tile's first cell and rounded once
// per tile; cell offsets within a tile are exact for uniform meshes.  NB models share
// every pair's multipole terms; model b's sums go to
// farpart[((b * nchunks + chunk) * NC + m) * S + s].
// Minimum waves per SIMD the far-field kernels are built for, by models per launch (1:
// the register count decides).  One model: six (at most 80 VGPRs; no scratch with the cube
// forms), measured faster than the five that 93 VGPRs allow and than seven or eight, whose
// spills reach the loops.  Two models in the forward and four in the adjoint: four waves
// (2.12 -> 1.96 and 1.35 -> 1.28 ms per model at 128x128x64 cells, 4096 sites); more waves
// for the others were no faster or spilled in the loops.  The single-model adjoint of other
// cell shapes holds four cells' multipole constants per lane and spilled 100-350 bytes per
// lane at six waves: four (99-128 VGPRs) halved it (config 1's 64x64x32 layered mesh
// 0.82 -> 0.42 ms, the vertical gradient's slab mesh 32.2 -> 21.1 ms per forward + adjoint;
// three waves, no scratch at all, measured within 3 % either way).  Four misfit vectors on
// a general mesh spilled ~300 bytes per lane at four waves: two (~190 VGPRs, no scratch) took
// config 1's batch of four from 1.21 to 1.03 ms (on config 2's uniform mesh two waves were
// 4.5 % slower than four, which spill 40 bytes).  The z column alone on a uniform mesh (the
// vertical gradient's slabs) needs fewer registers: eight waves forward (59 VGPRs; 3.88 ->
// 3.68 ms at 128x128 columns of 65 slabs, 4096 sites), six adjoint (80; 3.91 -> 3.87 ms).
constexpr int far_fwd_waves(int nb, int mesh, int mask) {
  return nb == 1 ? (mask == kMaskZcol && mesh == kMeshUniform ? 8 : 6) : nb == 2 ? 4 : 1;
}
constexpr int far_adj_waves(int nb, int mesh, int mask) {
  return nb == 1 ? (mesh == kMeshCubes || (mask == kMaskZcol && mesh == kMeshUniform) ? 6 : 4)
         : nb == 4 ? (mesh == kMeshGeneral ? 2 : 4)
                   : 1;
}
template <int MASK, int MESH, int NB>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(far_fwd_waves(NB, MESH, MASK)))) void k_far_forward(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const int* __restrict__ order, const int* __restrict__ box, float q1sq, FarGeom g,
    const float* __restrict__ shape, const float* __restrict__ wt, MpConst uc, long ntiles,
    long tiles_per_chunk, double* __restrict__ farpart) {
  constexpr int NC = Ncomp<MASK>::value;
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = order[li < S ? li : S - 1];
  const double px = sx[s], py = sy[s], pz = sz[s];
  int bx[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) bx[q] = box[s * 6 + q];
  const long t0 = (long)blockIdx.y * tiles_per_chunk;
  const long t1 = t0 + tiles_per_chunk < ntiles ? t0 + tiles_per_chunk : ntiles;
  double acc64[NB][NC];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int m = 0; m < NC; ++m) acc64[b][m] = 0.0;
  for (long t = t0; t < t1; ++t) {
    const int tk = (int)(t % g.ntz);
    const long tt = t / g.ntz;
    const int tj = (int)(tt % g.nty);
    const int ti = (int)(tt / g.nty);
    const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * kTile;
    const int ni = min(kTile, g.nx - i0), nj = min(kTile, g.ny - j0), nk = min(kTile, g.nzf - k0);
    const float Xt = (float)(px - g.xcen[i0]), Yt = (float)(py - g.ycen[j0]),
                Zt = (float)(pz - g.zcen[k0]);
    // near box against the mesh layers the tile's far layers span
    const bool hit = !(bx[1] < i0 || bx[0] > i0 + ni - 1 || bx[3] < j0 || bx[2] > j0 + nj - 1 ||
                       bx[5] < mesh_layer<MESH>(g, k0) || bx[4] > mesh_layer<MESH>(g, k0 + nk - 1));
    // some cell of the tile closer than the second radius, q1 half-diagonals of the tile's
    // largest cell: keep the hexadecapole (the offsets to the tile's cell centres span
    // [Xt - extent, Xt] along each axis)
    const float dx = fmaxf(0.0f, fmaxf(-Xt, Xt - g.xoff[i0 + ni - 1]));
    const float dy = fmaxf(0.0f, fmaxf(-Yt, Yt - g.yoff[j0 + nj - 1]));
    const float dz = fmaxf(0.0f, fmaxf(-Zt, Zt - g.zoff[k0 + nk - 1]));
    const float r1sq = q1sq * (g.hx2[ti] + g.hy2[tj] + g.hz2[tk]);
    const bool need = fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < r1sq;
    if (__any(hit))
      far_tile<MASK, MESH, true, true, NB>(g, shape, wt, uc, i0, j0, k0, ni, nj, nk, Xt, Yt, Zt,
                                           bx, acc64);
    else if (__any(need))
      far_tile<MASK, MESH, true, false, NB>(g, shape, wt, uc, i0, j0, k0, ni, nj, nk, Xt, Yt,
                                            Zt, bx, acc64);
    else if constexpr (MESH == kMeshCubes)
      far_tile_mono<MASK, NB>(g, wt, i0, j0, k0, ni, nj, nk, Xt, Yt, Zt, acc64);
    else
      far_tile<MASK, MESH, false, false, NB>(g, shape, wt, uc, i0, j0, k0, ni, nj, nk, Xt, Yt,
                                             Zt, bx, acc64);
  }
  if (li < S) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int m = 0; m < NC; ++m)
        farpart[(((long)b * gridDim.y + blockIdx.y) * NC + m) * S + s] = acc64[b][m];
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
  // block sums: within each wave by lane shuffles, then the four wave sums in fixed order
  __shared__ double red[NC][kBlock / 64];
  double sc[NC];
  comp_scales<MASK>(sc);
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    double v = acc[m];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[m][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int m = 0; m < NC; ++m)
      nearout[(long)s * NC + m] = ((red[m][0] + red[m][1]) + (red[m][2] + red[m][3])) * sc[m];
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

// Near plus far sums of one weight set, kept per site (component order of MASK).
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_mixed_forward_sums(const double* __restrict__ farpart,
                                                               int nchunks,
                                                               const double* __restrict__ nearout,
                                                               int S, double* __restrict__ sums) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
#pragma unroll
  for (int m = 0; m < NC; ++m) {
    double a = nearout[(long)s * NC + m];
    for (int c = 0; c < nchunks; ++c) a += farpart[((long)c * NC + m) * S + s];
    sums[(long)s * NC + m] = a;
  }
}

// Magnetization (OMPMagnetizationImp.cpp:25-117): B_i = scale sum_j T_ij M_j from the
// tensor sums U^j of the three weight sets Mx, My, Mz (sums[j][s][6], tensor order
// xx, yy, zz, xy, xz, yz), then the data transform.
__global__ __launch_bounds__(kBlock) void k_mag3_mixed_finish(const double* __restrict__ sums,
                                                              int S, double scale, int tf,
                                                              double tfbx, double tfby,
                                                              double tfbz,
                                                              double* __restrict__ out,
                                                              double* __restrict__ raw_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double u[3][7];
  for (int j = 0; j < 3; ++j) spread_components<kMaskTensor>(sums + ((long)j * S + s) * 6, u[j]);
  // u[.]: gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz
  double raw[3];
  raw[0] = scale * (u[0][1] + u[1][4] + u[2][5]);
  raw[1] = scale * (u[0][4] + u[1][2] + u[2][6]);
  raw[2] = scale * (u[0][5] + u[1][6] + u[2][3]);
  store_site(raw, 3, s, tf, tfbx, tfby, tfbz, out, raw_out);
}

// Magnetization adjoint as three tensor adjoints: dB_i / dM_j = scale T_ij, so the
// gradient of M_j is the tensor gradient of the raw tensor misfit R^j with
// R^j_(i, j) = scale r_i and zeros elsewhere (raw tensor order xx, xy, xz, yx, ...).
__global__ __launch_bounds__(kBlock) void k_mag3_tensor_misfit(const double* __restrict__ rm,
                                                               int S, double scale,
                                                               double* __restrict__ rmt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 27L * S) return;
  const int j = (int)(idx / (9L * S));
  const long rest = idx - (long)j * 9 * S;
  const int s = (int)(rest / 9);
  const int q = (int)(rest % 9);
  const int i = q / 3;
  rmt[idx] = (q % 3 == j) ? scale * rm[(long)s * 3 + i] : 0.0;
}

// Adjoint coefficients: coef (FP64, with the sign and G of each corner sum, for
// the near pass) and coef32 (FP32 physical values, for the far pass).
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_mixed_coef(const double* __restrict__ rm, int S,
                                                       int kind, FieldDir fd,
                                                       double* __restrict__ c64,
                                                       float* __restrict__ c32,
                                                       float* __restrict__ mono) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double c7[7], c[NC], sc[NC];
  site_coefficients(rm, s, kind, fd, c7);
  select_components<MASK>(c7, c);
  comp_scales<MASK>(sc);
  for (int m = 0; m < NC; ++m) {
    c64[(long)s * NC + m] = c[m] * sc[m];
    c32[(long)s * NC + m] = (float)c[m];
  }
  // monopole form of sum_m c_m T_m: the monopole tensor is trace-free, so the diagonal
  // coefficients may drop their mean t/3 (t = c_xx + c_yy + c_zz), and with s^3 = r^2 s^5
  //   sum_m c_m T_m = s^5 (Q(X, Y, Z) - c_gz Z r^2),
  //   Q = 3 ((c_xx - t/3) X^2 + (c_yy - t/3) Y^2 + (c_zz - t/3) Z^2 + c_xy XY + c_xz XZ
  //       + c_yz YZ);
  // stored: -c_gz, the six factors of Q, 0; scaled by kFixS5 for the bias of the raw
  // s^5 of the monopole path (rounded to FP32 per site, the scaling holds on average over
  // the sites a cell's sum runs over)
  float* o = mono + (long)s * 8;
  const double t3 = (c7[1] + c7[2] + c7[3]) / 3.0;
  o[0] = (float)(-c7[0] * kFixS5);
  for (int m = 1; m < 4; ++m) o[m] = (float)(3.0 * (c7[m] - t3) * kFixS5);
  for (int m = 4; m < 7; ++m) o[m] = (float)(3.0 * c7[m] * kFixS5);
  o[7] = 0.0f;
}

// Adjoint far field, one site's terms for a lane's kQ cells, added to acc.  HIT: the
// site's near box touches the tile, and the lane's cells inside it (nq) are moved far
// away and dropped (FP64 near pass); otherwise no cell needs a mask.
constexpr int kQ = 4;  // cells per lane of the far adjoint
template <int MASK, int MESH, int NB, bool HIT>
__device__ __forceinline__ void far_adjoint_site(
    const float* __restrict__ coef, const float4* mono4, int S, int s, float Xt, float Yt,
    float Zt, bool need, float ox, float oy, const float (&ozq)[kQ], const MpConst (&ccq)[kQ],
    const bool (&nq)[kQ], float (&acc)[NB][kQ]) {
  constexpr int NC = Ncomp<MASK>::value;
  constexpr bool QUAD = MESH != kMeshCubes;
  const float X = Xt - ox, Y = Yt - oy;
  float Zq[kQ], pq[NB][kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) Zq[q] = (HIT && nq[q]) ? 1.0e6f : Zt - ozq[q];
  if (MESH == kMeshCubes && !need) {
    // monopole with the site's regrouped, trace-free coefficients:
    // s^5 (Q - c_gz Z r^2) = s^5 (Qxy + Z (Lz - c_gz (X^2 + Y^2) + Z (3 c'_zz - c_gz Z)))
    const float XX = X * X, YY = Y * Y, XY = X * Y;
    const float XXYY = XX + YY;
    float Qxy[NB], L2[NB], ngz[NB], m3[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const float4 lo = mono4[2 * b], hi = mono4[2 * b + 1];  // mc[0..3], mc[4..7]
      Qxy[b] = fmaf(lo.y, XX, fmaf(lo.z, YY, hi.x * XY));
      L2[b] = fmaf(lo.x, XXYY, fmaf(hi.y, X, hi.z * Y));
      ngz[b] = lo.x;
      m3[b] = lo.w;
    }
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      const float Z = Zq[q];
      const float r2 = fmaf(Z, Z, XXYY);
      const float sv = __builtin_amdgcn_rsqf(r2);  // bias removed in the coefficients
      const float s2 = sv * sv;
      const float s5 = (s2 * sv) * s2;
#pragma unroll
      for (int b = 0; b < NB; ++b)
        pq[b][q] = fmaf(Z, fmaf(Z, fmaf(ngz[b], Z, m3[b]), L2[b]), Qxy[b]) * s5;
    }
  } else {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      float T[NC];
      if constexpr (MESH == kMeshCubes) {
        if (need)
          multipole_cube_f32<MASK>(X, Y, Zq[q], 0.25f * ccq[q].a4x, T);
        else
          multipole_f32<MASK, QUAD, false>(X, Y, Zq[q], ccq[q], T);
      } else if (need) {
        multipole_f32<MASK, QUAD, true>(X, Y, Zq[q], ccq[q], T);
      } else {
        multipole_f32<MASK, QUAD, false>(X, Y, Zq[q], ccq[q], T);
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const float* cs = coef + ((long)b * S + s) * NC;
        float p = 0.0f;
#pragma unroll
        for (int m = 0; m < NC; ++m) p = fmaf(cs[m], T[m], p);
        pq[b][q] = p;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < kQ; ++q) acc[b][q] += (HIT && nq[q]) ? 0.0f : pq[b][q];
}

// Adjoint far field: one lane per kQ cells (k, k + 4, k + 8, ...) of a 4x4x(4 kQ) tile;
// the wave walks a chunk of sites (wave-uniform) and accumulates
// sum_m coef_m(s) T_m(s, c) in FP32, flushed to FP64 every 8 sites; cells in the site's
// near box are skipped (FP64 near pass).  Site offsets are formed in FP64 against the
// tile's first cell and rounded once; a site whose second radius misses every cell of
// the tile keeps only the monopole (and quadrupole) terms, a wave-uniform branch.  These
// per-site values, whether the site's near box touches the tile and the site's monopole
// coefficients are the same for every lane, so each lane forms or loads them for one site
// of a run of 64 and the wave reads them back from LDS (no scalar-cache misses per site).  NB misfit vectors share the kernel
// terms T_m; model b's coefficients sit at coef + b S NC and mono + b S 8, its partial
// sums at gpart[(b nchunks + chunk) ncell + c].
template <int MASK, int MESH, int NB>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(far_adj_waves(NB, MESH, MASK)))) void k_far_adjoint(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const float* __restrict__ coef, const float* __restrict__ mono, const int* __restrict__ box,
    float q1sq, int S, int nchunks, const int* __restrict__ ranges, FarGeom g,
    const float* __restrict__ table, MpConst uc, long ntiles, double* __restrict__ gpart) {
  constexpr int TW = TableWidth<MESH>::value;
  constexpr int KT = kQ * kTile;  // tile depth in k
  // per-site values of the wave's run of 64 sites: offsets and flags, and the NB sets of
  // monopole coefficients (two float4 each)
  __shared__ float4 sitev[kBlock];
  __shared__ float4 monov[kBlock * 2 * NB];
  const long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  const int run = threadIdx.x - lane;
  if (t >= ntiles) return;
  const int ntz = (g.nzf + KT - 1) / KT;
  const int tk = (int)(t % ntz);
  const long tt = t / ntz;
  const int tj = (int)(tt % g.nty);
  const int ti = (int)(tt / g.nty);
  const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * KT;
  const int ni = min(kTile, g.nx - i0), nj = min(kTile, g.ny - j0), nk = min(KT, g.nzf - k0);
  // mesh layers spanned by the tile's far layers (near-box tests)
  const int kml = mesh_layer<MESH>(g, k0), kmh = mesh_layer<MESH>(g, k0 + nk - 1);
  const int ii = lane >> 4, jj = (lane >> 2) & 3, kk = lane & 3;
  const bool vij = ii < ni && jj < nj;
  const int i = i0 + (vij ? ii : 0), j = j0 + (vij ? jj : 0);
  const double x0 = g.xcen[i0], y0 = g.ycen[j0], z0 = g.zcen[k0];
  const float ox = (float)(g.xcen[i] - x0), oy = (float)(g.ycen[j] - y0);
  // extent of the tile's cell centres from its first cell, per axis
  const float ex = (float)(g.xcen[i0 + ni - 1] - x0), ey = (float)(g.ycen[j0 + nj - 1] - y0),
              ez = (float)(g.zcen[k0 + nk - 1] - z0);
  // the second radius: q1 half-diagonals of the tile's largest cell
  const float r1sq = q1sq * (g.hx2[ti] + g.hy2[tj] + g.hz2_16[tk]);
  bool vq[kQ];
  int kq[kQ];  // mesh layer of each of the lane's cells
  float ozq[kQ];
  MpConst ccq[kQ];
  // cell index of the lane's cell q (formed again for the output rather than held through
  // the site loop)
  const long cij = ((long)i * g.ny + j) * g.nz;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    vq[q] = vij && kk + q * kTile < nk;
    const int f = k0 + (vq[q] ? kk + q * kTile : 0);  // far layer
    kq[q] = mesh_layer<MESH>(g, f);
    ozq[q] = (float)(g.zcen[f] - z0);
    ccq[q] = cell_const<MESH>(table, cij + kq[q], uc);
  }
  // this chunk's sites: a range of the planned partition (equal estimated cost for every
  // chunk of the workgroup's four tiles)
  const int s0 = ranges[blockIdx.x * (nchunks + 1) + blockIdx.y];
  const int s1 = ranges[blockIdx.x * (nchunks + 1) + blockIdx.y + 1];
  double a64[NB][kQ];
  float acc[NB][kQ];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      a64[b][q] = 0.0;
      acc[b][q] = 0.0f;
    }
  int nacc = 0;
  for (int sb = s0; sb < s1; sb += 64) {
    const int nsb = min(64, s1 - sb);
    {
      // this lane's site of the run: offsets from the tile, the second-radius test (bit
      // 0), whether its near box touches the tile (bit 1), its monopole coefficients
      const int sl = sb + (lane < nsb ? lane : 0);
      const float Xl = (float)(sx[sl] - x0), Yl = (float)(sy[sl] - y0),
                  Zl = (float)(sz[sl] - z0);
      const float dx = fmaxf(0.0f, fmaxf(-Xl, Xl - ex)), dy = fmaxf(0.0f, fmaxf(-Yl, Yl - ey)),
                  dz = fmaxf(0.0f, fmaxf(-Zl, Zl - ez));
      const bool nl = fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < r1sq;
      const int* bl = box + sl * 6;
      const bool hl = !(bl[1] < i0 || bl[0] > i0 + ni - 1 || bl[3] < j0 ||
                        bl[2] > j0 + nj - 1 || bl[5] < kml || bl[4] > kmh);
      sitev[threadIdx.x] = make_float4(Xl, Yl, Zl, (float)((nl ? 1 : 0) | (hl ? 2 : 0)));
      if (MESH == kMeshCubes) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const float4* m4 = reinterpret_cast<const float4*>(mono + ((long)b * S + sl) * 8);
          monov[threadIdx.x * 2 * NB + 2 * b] = m4[0];
          monov[threadIdx.x * 2 * NB + 2 * b + 1] = m4[1];
        }
      }
    }
    // the wave's own LDS writes are complete before its reads below
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int js = 0; js < nsb; ++js) {
      const int s = sb + js;
      const float4 v = sitev[run + js];
      const int flags = __builtin_amdgcn_readfirstlane((int)v.w);
      const bool need = (flags & 1) != 0;
      const float4* m4 = monov + (run + js) * 2 * NB;
      bool nq[kQ];
      if (flags & 2) {
        const int* bx = box + s * 6;
        const bool inxy = i >= bx[0] && i <= bx[1] && j >= bx[2] && j <= bx[3];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
          // for one misfit vector on cubes formed here instead of held through the loop
          // (spilled there at six waves); past the mesh (vq false) the test decides nothing
          // that is stored
          const int km = MESH == kMeshCubes && NB == 1 ? k0 + kk + q * kTile : kq[q];
          nq[q] = inxy && km >= bx[4] && km <= bx[5];
        }
        far_adjoint_site<MASK, MESH, NB, true>(coef, m4, S, s, v.x, v.y, v.z, need, ox, oy, ozq,
                                               ccq, nq, acc);
      } else {
#pragma unroll
        for (int q = 0; q < kQ; ++q) nq[q] = false;
        far_adjoint_site<MASK, MESH, NB, false>(coef, m4, S, s, v.x, v.y, v.z, need, ox, oy,
                                                ozq, ccq, nq, acc);
      }
      if (++nacc == 8) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
          for (int q = 0; q < kQ; ++q) {
            a64[b][q] += (double)acc[b][q];
            acc[b][q] = 0.0f;
          }
        nacc = 0;
      }
    }
    // every read of this run is done before the next run overwrites the slots
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  const long ncell = (long)g.nx * g.ny * g.nz;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int km = MESH == kMeshCubes && NB == 1 ? k0 + (vq[q] ? kk + q * kTile : 0) : kq[q];
    const long cq = ((long)i * g.ny + j) * g.nz + km;
    const double gv = 2.0 * (double)table[cq * TW];  // G V
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      a64[b][q] += (double)acc[b][q];
      if (vq[q]) gpart[((long)b * gridDim.y + blockIdx.y) * ncell + cq] = gv * a64[b][q];
    }
  }
}

// Adjoint near field, pass 1: one block per site.  H(corner) = sum_m coef_m F_m
// over the corners of the site's near box (FP64 closed form, each corner once),
// then v(cell) = 2 sum sigma H over the cell's corners for every cell of the box.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_near_adjoint_site(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    const int* __restrict__ box, const double* __restrict__ coef,
    const double* __restrict__ Xc, const double* __restrict__ Yc,
    const double* __restrict__ Zc, const long* __restrict__ corner_off,
    const long* __restrict__ cell_off, double* __restrict__ hbuf,
    double* __restrict__ vbuf) {
  constexpr int NC = Ncomp<MASK>::value;
  const int s = blockIdx.x;
  const int i0 = box[s * 6 + 0], i1 = box[s * 6 + 1];
  const int j0 = box[s * 6 + 2], j1 = box[s * 6 + 3];
  const int k0 = box[s * 6 + 4], k1 = box[s * 6 + 5];
  if (i0 > i1 || j0 > j1 || k0 > k1) return;
  const double px = sx[s], py = sy[s], pz = sz[s];
  double cs[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) cs[m] = coef[(long)s * NC + m];
  const int ci = i1 - i0 + 2, cj = j1 - j0 + 2, ck = k1 - k0 + 2;
  double* H = hbuf + corner_off[s];
  for (int q = threadIdx.x; q < ci * cj * ck; q += blockDim.x) {
    const int K = k0 + q % ck;
    const int J = j0 + (q / ck) % cj;
    const int I = i0 + q / (ck * cj);
    double t[NC];
    corner_terms_f64<MASK>(px - Xc[I], py - Yc[J], pz - Zc[K], t);
    double h = 0.0;
#pragma unroll
    for (int m = 0; m < NC; ++m) h = fma(cs[m], t[m], h);
    H[q] = h;
  }
  __threadfence_block();
  __syncthreads();
  const int ni = ci - 1, nj = cj - 1, nk = ck - 1;
  double* V = vbuf + cell_off[s];
  for (int q = threadIdx.x; q < ni * nj * nk; q += blockDim.x) {
    const int kk = q % nk;
    const int jj = (q / nk) % nj;
    const int ii = q / (nk * nj);
    double a = 0.0;
#pragma unroll
    for (int da = 0; da < 2; ++da)
#pragma unroll
      for (int db = 0; db < 2; ++db)
#pragma unroll
        for (int dc = 0; dc < 2; ++dc) {
          const double h = H[((ii + da) * cj + (jj + db)) * ck + (kk + dc)];
          a += ((da + db + dc) & 1) ? h : -h;
        }
    V[q] = 2.0 * a;
  }
}

// Adjoint near field, pass 2: one 64-lane block per listed column (i, j); each cell sums
// the values of the sites whose box holds it, in ascending site order.  The column's sites
// (their box's k range and the offset of the column in their value buffer) are staged in
// LDS 64 at a time, so the value loads of a cell do not wait on one another.
__global__ __launch_bounds__(64) void k_near_adjoint_gather(
    const int* __restrict__ col_ptr, const int* __restrict__ col_sites,
    const int* __restrict__ col_ids, const int* __restrict__ col_kmax,
    const int* __restrict__ box, const long* __restrict__ cell_off,
    const double* __restrict__ vbuf, int ny, int nz, double* __restrict__ gnear) {
  __shared__ long soff[64];  // vbuf index of (i, j, k = 0) of the site's box run
  __shared__ int sk0[64], sk1[64];
  const int col = col_ids[blockIdx.x];
  const int i = col / ny, j = col % ny;
  const int p0 = col_ptr[blockIdx.x], p1 = col_ptr[blockIdx.x + 1];
  const int kmax = col_kmax[blockIdx.x];
  for (int kb = 0; kb <= kmax; kb += 64) {
    const int k = kb + (int)threadIdx.x;
    double acc = 0.0;
    for (int pb = p0; pb < p1; pb += 64) {
      const int np = min(64, p1 - pb);
      __syncthreads();  // the previous run of sites has been read
      if ((int)threadIdx.x < np) {
        const int s = col_sites[pb + threadIdx.x];
        const int* b = box + s * 6;
        const int nj = b[3] - b[2] + 1, nk = b[5] - b[4] + 1;
        soff[threadIdx.x] = cell_off[s] + ((long)(i - b[0]) * nj + (j - b[2])) * nk - b[4];
        sk0[threadIdx.x] = b[4];
        sk1[threadIdx.x] = b[5];
      }
      __syncthreads();
      if (k <= kmax)
        for (int q = 0; q < np; ++q)
          if (k >= sk0[q] && k <= sk1[q]) acc += vbuf[soff[q] + k];
    }
    if (k <= kmax) gnear[((long)i * ny + j) * nz + k] = acc;
  }
}

// grad = near + far chunk sums per cell; cells of layers outside a far-layer list (listed
// non-null, listed[k] == 0) get zero, their far sums never having been formed
__global__ __launch_bounds__(kBlock) void k_adjoint_mixed_finish(const double* __restrict__ gpart,
                                                                 int nchunks,
                                                                 const double* __restrict__ gnear,
                                                                 long n,
                                                                 const int* __restrict__ listed,
                                                                 int nz,
                                                                 double* __restrict__ grad) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (listed && !listed[c % nz]) {
    grad[c] = 0.0;
    return;
  }
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
  g.xoff = ctx->near.xoff.as<const float>();
  g.yoff = ctx->near.yoff.as<const float>();
  g.zoff = ctx->near.zoff.as<const float>();
  g.hx2 = ctx->near.hx2.as<const float>();
  g.hy2 = ctx->near.hy2.as<const float>();
  g.hz2 = ctx->near.hz2.as<const float>();
  g.hz2_16 = ctx->near.hz2_16.as<const float>();
  g.kmap = ctx->near.far_subset ? ctx->near.kmap.as<const int>() : nullptr;
  g.nx = (int)ctx->nx;
  g.ny = (int)ctx->ny;
  g.nz = (int)ctx->nz;
  g.nzf = ctx->near.far_subset ? ctx->near.nzf : g.nz;
  g.ntx = (g.nx + kTile - 1) / kTile;
  g.nty = (g.ny + kTile - 1) / kTile;
  g.ntz = (g.nzf + kTile - 1) / kTile;
  return g;
}

// Square of the second radius in tile half-diagonals (q1^2; a tile's cells closer than q1
// of its largest cell's half-diagonals keep the hexadecapole); every cell when q1 is 0.
float second_radius_q2(const jgm_context* ctx) {
  if (!(ctx->mixed_q1 > 0.0)) return FLT_MAX;
  return (float)(ctx->mixed_q1 * ctx->mixed_q1);
}

MpConst uniform_const(const jgm_context* ctx) {
  if (ctx->nx == 0 || ctx->ny == 0 || ctx->nz == 0) return MpConst{};
  return mp_const(ctx->xs[0], ctx->ys[0], ctx->near.far_subset ? ctx->near.far_dz : ctx->zs[0]);
}

// Sites sorted by the Morton code of their (x, y) position in the survey's bounding
// box, so that the 64 lanes of a far-field wave cover a compact patch of sites.
std::vector<int> morton_order(const jgm_context* ctx) {
  const size_t S = ctx->S;
  std::vector<int> order(std::max<size_t>(1, S), 0);
  if (S == 0) return order;
  double x0 = ctx->hsx[0], x1 = x0, y0 = ctx->hsy[0], y1 = y0;
  for (size_t s = 0; s < S; ++s) {
    x0 = std::min(x0, ctx->hsx[s]); x1 = std::max(x1, ctx->hsx[s]);
    y0 = std::min(y0, ctx->hsy[s]); y1 = std::max(y1, ctx->hsy[s]);
  }
  const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
  auto spread = [](uint32_t v) {  // 16 bits -> even bit positions
    v &= 0xFFFF;
    v = (v | (v << 8)) & 0x00FF00FF;
    v = (v | (v << 4)) & 0x0F0F0F0F;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
  };
  std::vector<uint32_t> code(S);
  for (size_t s = 0; s < S; ++s) {
    const uint32_t qx = (uint32_t)((ctx->hsx[s] - x0) * sx), qy = (uint32_t)((ctx->hsy[s] - y0) * sy);
    code[s] = spread(qx) | (spread(qy) << 1);
    order[s] = (int)s;
  }
  std::stable_sort(order.begin(), order.begin() + S,
                   [&](int a, int b) { return code[a] < code[b]; });
  return order;
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

const char* far_name(int mask, bool adjoint) {
  if (adjoint)
    return mask == kMaskScalar ? "far_adjoint_fp32_gz"
           : mask == kMaskAll  ? "far_adjoint_fp32_gz_ftg"
           : mask == kMaskZcol ? "far_adjoint_fp32_ftg_zcol"
                               : "far_adjoint_fp32_ftg";
  return mask == kMaskScalar ? "far_forward_fp32_gz"
         : mask == kMaskAll  ? "far_forward_fp32_gz_ftg"
         : mask == kMaskZcol ? "far_forward_fp32_ftg_zcol"
                             : "far_forward_fp32_ftg";
}

// The far adjoint's work partition: the sites of each workgroup of four adjoint tiles are cut
// into nchunks ranges of equal estimated cost, each site's cost being, per tile, its path as
// k_far_adjoint takes it (masked near-box tile, hexadecapole within the second radius,
// monopole beyond) times the tile's cells, in integer units (cubes 2 / 12 / 13, other
// shapes 10 / 22 / 25: the kernels' instruction counts per pair).  Block tb writes
// ranges[tb (nchunks + 1) + j], the first site at which the running cost reaches j / nchunks
// of the block's total, from an exact integer prefix sum (so the same for every launch).
template <int MESH>
__device__ __forceinline__ long long adjoint_site_cost(const double* __restrict__ sx,
                                                       const double* __restrict__ sy,
                                                       const double* __restrict__ sz,
                                                       const int* __restrict__ box, float q1sq,
                                                       const FarGeom& g, long ntiles, long tb,
                                                       int s) {
  constexpr int KT = kQ * kTile;
  const int ntz = (g.nzf + KT - 1) / KT;
  const long long cm = MESH == kMeshCubes ? 2 : 10, ch = MESH == kMeshCubes ? 12 : 22,
                  ck = MESH == kMeshCubes ? 13 : 25;
  long long c = 0;
  for (long t = 4 * tb; t < 4 * tb + 4 && t < ntiles; ++t) {
    // the tile's geometry and the site's tests as in k_far_adjoint
    const int tk = (int)(t % ntz);
    const long tt = t / ntz;
    const int tj = (int)(tt % g.nty), ti = (int)(tt / g.nty);
    const int i0 = ti * kTile, j0 = tj * kTile, k0 = tk * KT;
    const int ni = min(kTile, g.nx - i0), nj = min(kTile, g.ny - j0), nk = min(KT, g.nzf - k0);
    const double x0 = g.xcen[i0], y0 = g.ycen[j0], z0 = g.zcen[k0];
    const float ex = (float)(g.xcen[i0 + ni - 1] - x0), ey = (float)(g.ycen[j0 + nj - 1] - y0),
                ez = (float)(g.zcen[k0 + nk - 1] - z0);
    const float r1sq = q1sq * (g.hx2[ti] + g.hy2[tj] + g.hz2_16[tk]);
    const float Xl = (float)(sx[s] - x0), Yl = (float)(sy[s] - y0), Zl = (float)(sz[s] - z0);
    const float dx = fmaxf(0.0f, fmaxf(-Xl, Xl - ex)), dy = fmaxf(0.0f, fmaxf(-Yl, Yl - ey)),
                dz = fmaxf(0.0f, fmaxf(-Zl, Zl - ez));
    const bool need = fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < r1sq;
    const int* bl = box + s * 6;
    const bool hit = !(bl[1] < i0 || bl[0] > i0 + ni - 1 || bl[3] < j0 || bl[2] > j0 + nj - 1 ||
                       bl[5] < mesh_layer<MESH>(g, k0) || bl[4] > mesh_layer<MESH>(g, k0 + nk - 1));
    c += (hit ? ck : need ? ch : cm) * (long long)(ni * nj * nk);
  }
  return c;
}

// Split points of a block's sequence of n costs cost(i) >= 0 (one workgroup of kBlock threads):
// out[j] = the first index at which the running cost reaches j / nchunks of the total, from an
// exact integer prefix sum; out[0] = 0, out[nchunks] = n.
template <class F>
__device__ void split_by_cost(long n, int nchunks, F cost, long long* scan, int* out) {
  const int tid = threadIdx.x;
  long long part = 0;
  for (long i = tid; i < n; i += kBlock) part += cost(i);
  scan[tid] = part;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (tid < w) scan[tid] += scan[tid + w];
    __syncthreads();
  }
  const long long total = scan[0];
  __syncthreads();
  if (tid == 0) {
    out[0] = 0;
    out[nchunks] = (int)n;
  }
  // running prefix in runs of kBlock (inclusive scan in LDS); split j goes to the index whose
  // cost interval (run + excl, run + incl] holds j total / nchunks
  long long run = 0;
  for (long i0 = 0; i0 < n; i0 += kBlock) {
    const long i = i0 + tid;
    const long long c = i < n ? cost(i) : 0;
    scan[tid] = c;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      const long long v = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    const long long incl = run + scan[tid], excl = incl - c;
    if (c > 0 && total > 0) {
      const long long jlo = excl * nchunks / total + 1, jhi = incl * nchunks / total;
      for (long long j = jlo; j <= jhi && j < nchunks; ++j)
        if (j >= 1) out[j] = (int)(i + 1);
    }
    run += scan[kBlock - 1];
    __syncthreads();
  }
}

template <int MESH>
__global__ __launch_bounds__(kBlock) void k_far_adjoint_plan(
    const double* __restrict__ sx, const double* __restrict__ sy, const double* __restrict__ sz,
    int S, const int* __restrict__ box, float q1sq, FarGeom g, long ntiles, int nchunks,
    int* __restrict__ ranges) {
  __shared__ long long scan[kBlock];
  const long tb = blockIdx.x;
  split_by_cost(
      S, nchunks,
      [&](long s) {
        return adjoint_site_cost<MESH>(sx, sy, sz, box, q1sq, g, ntiles, tb, (int)s);
      },
      scan, ranges + tb * (nchunks + 1));
}

// Makes the far adjoint's partition if the geometry, the second radius or the wave target
// changed (stream-ordered before the adjoint that reads it; no host work beyond the launch).
// At 128x128x64 cells and 4096 sites, 12 planned chunks (49152 waves) ran the far adjoint in
// 1.83 ms against 1.84 ms for 24 equal ones, with half the FP64 partial sums (0.10 GB).
template <int MESH>
hipError_t ensure_far_plan(jgm_context* ctx) {
  NearBoxes& nb = ctx->near;
  if (nb.plan_valid && nb.plan_q1 == ctx->mixed_q1) return hipSuccess;
  const int S = (int)ctx->S;
  const FarGeom g = far_geom(ctx);
  const long ntiles = (long)g.ntx * g.nty * ((g.nzf + kQ * kTile - 1) / (kQ * kTile));
  const long nblocks = (ntiles + 3) / 4;
  long C = (ctx->far_adj_waves + 4 * nblocks - 1) / std::max(1L, 4 * nblocks);
  C = std::max<long>(1, std::min<long>(C, std::max(1, S)));
  hipError_t e = nb.adj_ranges.ensure(sizeof(int) * (size_t)std::max(1L, nblocks) * (C + 1));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_far_adjoint_plan<MESH>), dim3((unsigned)nblocks), dim3(kBlock), 0,
                     ctx->stream, ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                     ctx->d_sz.as<const double>(), S, nb.box.as<const int>(),
                     second_radius_q2(ctx), g, ntiles, (int)C, nb.adj_ranges.as<int>());
  if ((e = hipGetLastError()) != hipSuccess) return e;
  nb.adj_chunks = (int)C;
  nb.plan_q1 = ctx->mixed_q1;
  nb.plan_valid = true;
  return hipSuccess;
}

// Unit-density cell table of the mesh (G V and, for general meshes, the shape
// constants), built once per geometry; the adjoint's weights and the forward's shapes.
template <int MESH>
hipError_t ensure_shape_table(jgm_context* ctx) {
  NearBoxes& nb = ctx->near;
  if (nb.adj_table_valid) return hipSuccess;
  const FarGeom g = far_geom(ctx);
  const long N = (long)ncells(ctx);
  hipError_t e = nb.adj_table.ensure(sizeof(float) * (size_t)N * TableWidth<MESH>::value);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_cell_table<MESH>), dim3(blocks_for(N)), dim3(kBlock), 0, ctx->stream,
                     (const double*)nullptr, ctx->d_xsz.as<const double>(),
                     ctx->d_ysz.as<const double>(), ctx->d_zsz.as<const double>(), g.nx, g.ny,
                     g.nz, nb.adj_table.as<float>());
  nb.adj_table_valid = true;
  return hipGetLastError();
}

// nb in {1, 2, 4}, or 3 with the tensor mask (the magnetization kind's Mx, My, Mz)
template <int MASK, int MESH, class... Args>
void launch_far_forward(int nb, dim3 grid, hipStream_t st, Args... args) {
  switch (nb) {
    case 4:
      hipLaunchKernelGGL((k_far_forward<MASK, MESH, 4>), grid, dim3(kBlock), 0, st, args...);
      break;
    case 3:
      if constexpr (MASK == kMaskTensor)
        hipLaunchKernelGGL((k_far_forward<MASK, MESH, 3>), grid, dim3(kBlock), 0, st, args...);
      break;
    case 2:
      hipLaunchKernelGGL((k_far_forward<MASK, MESH, 2>), grid, dim3(kBlock), 0, st, args...);
      break;
    default:
      hipLaunchKernelGGL((k_far_forward<MASK, MESH, 1>), grid, dim3(kBlock), 0, st, args...);
      break;
  }
}

// Forward of nb models (model b at models + b * model_ld; data at out + b * out_ld, raw
// data at raw_out + b * raw_ld when given): one far-field pass for all of them, the FP64
// near field and the epilogue per model.  With sums non-null the near plus far sums of
// weight set b go to sums + b S NC and no data are formed (the magnetization kind
// combines its three sets).
template <int MASK, int MESH>
hipError_t forward_mixed_impl(jgm_context* ctx, int kind, int nbm, const double* models,
                              size_t model_ld, double* out, size_t out_ld, double* raw_out,
                              size_t raw_ld, double* sums) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const FarGeom g = far_geom(ctx);
  const long ntiles = (long)g.ntx * g.nty * g.ntz;
  hipError_t e;
  int nchunks = 0;
  if ((e = ctx->nearpart.ensure(sizeof(double) * (size_t)nbm * S * NC)) != hipSuccess) return e;
  if (N > 0) {
    const int site_blocks = (S + kBlock - 1) / kBlock;
    const long waves_per_chunk = (long)site_blocks * (kBlock / 64);
    long want = (ctx->far_fwd_waves + waves_per_chunk - 1) / waves_per_chunk;
    want = std::max<long>(1, std::min<long>(want, ntiles));
    const long per = (ntiles + want - 1) / want;
    nchunks = (int)((ntiles + per - 1) / per);
    if ((e = ctx->weights.ensure(sizeof(float) * (size_t)N * nbm)) != hipSuccess) return e;
    if ((e = ctx->farpart.ensure(sizeof(double) * (size_t)nbm * nchunks * NC * S)) != hipSuccess)
      return e;
    if constexpr (MESH == kMeshGeneral) {
      if ((e = ensure_shape_table<MESH>(ctx)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_cell_weights, dim3(blocks_for(N)), dim3(kBlock), 0, st, models,
                       (long)model_ld, nbm, ctx->d_xsz.as<const double>(),
                       ctx->d_ysz.as<const double>(), ctx->d_zsz.as<const double>(), g.nx, g.ny,
                       g.nz, ctx->weights.as<float>());
    {
      ProfScope ps(ctx, far_name(MASK, false));
      launch_far_forward<MASK, MESH>(nbm, dim3(site_blocks, nchunks), st,
                                     ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                                     ctx->d_sz.as<const double>(), S,
                                     ctx->near.site_order.as<const int>(),
                                     ctx->near.box.as<const int>(), second_radius_q2(ctx), g,
                                     ctx->near.adj_table.as<const float>(),
                                     ctx->weights.as<const float>(), uniform_const(ctx), ntiles,
                                     per, ctx->farpart.as<double>());
    }
    {
      ProfScope ps(ctx, "near_forward_fp64");
      for (int b = 0; b < nbm; ++b)
        hipLaunchKernelGGL((k_near_forward<MASK>), dim3(S), dim3(kBlock), 0, st,
                           ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),
                           ctx->d_sz.as<double>(), ctx->near.box.as<const int>(),
                           ctx->d_Xc.as<double>(), ctx->d_Yc.as<double>(),
                           ctx->d_Zc.as<double>(), models + b * model_ld, g.nx, g.ny, g.nz,
                           ctx->nearpart.as<double>() + (size_t)b * S * NC);
    }
  } else {
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for((long)nbm * S * NC)), dim3(kBlock), 0, st,
                       ctx->nearpart.as<double>(), (size_t)nbm * S * NC);
  }
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  if (nchunks > 0 && (e = ctx->partsum.ensure(sizeof(double) * (size_t)NC * S)) != hipSuccess)
    return e;
  for (int b = 0; b < nbm; ++b) {
    if (nchunks > 0 &&
        (e = launch_reduce_chunks(ctx, ctx->farpart.as<const double>() + (size_t)b * nchunks * NC * S,
                                  nchunks, (long)NC * S, ctx->partsum.as<double>())) != hipSuccess)
      return e;
    if (sums)
      hipLaunchKernelGGL((k_mixed_forward_sums<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st,
                         ctx->partsum.as<const double>(), nchunks > 0 ? 1 : 0,
                         ctx->nearpart.as<const double>() + (size_t)b * S * NC, S,
                         sums + (size_t)b * S * NC);
    else
      hipLaunchKernelGGL((k_mixed_forward_finish<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0,
                         st, ctx->partsum.as<const double>(), nchunks > 0 ? 1 : 0,
                         ctx->nearpart.as<const double>() + (size_t)b * S * NC, S,
                         ctx->bgT.as<const double>(), models + b * model_ld + N, L, kind,
                         field_dir(ctx), ctx->transform, tf_bx, tf_by, tf_bz, out + b * out_ld,
                         raw_out ? raw_out + b * raw_ld : nullptr);
  }
  return hipGetLastError();
}

// nbm in {1, 2, 4}, or 3 with the tensor mask (the magnetization kind)
template <int MASK, int MESH, class... Args>
void launch_far_adjoint(int nbm, dim3 grid, hipStream_t st, Args... args) {
  switch (nbm) {
    case 4:
      hipLaunchKernelGGL((k_far_adjoint<MASK, MESH, 4>), grid, dim3(kBlock), 0, st, args...);
      break;
    case 3:
      if constexpr (MASK == kMaskTensor)
        hipLaunchKernelGGL((k_far_adjoint<MASK, MESH, 3>), grid, dim3(kBlock), 0, st, args...);
      break;
    case 2:
      hipLaunchKernelGGL((k_far_adjoint<MASK, MESH, 2>), grid, dim3(kBlock), 0, st, args...);
      break;
    default:
      hipLaunchKernelGGL((k_far_adjoint<MASK, MESH, 1>), grid, dim3(kBlock), 0, st, args...);
  }
}

// Gradients of nbm misfit vectors rm + b rm_ld into grad + b grad_ld; with_bg false: no
// background-layer entries after the cells (magnetization kind).
template <int MASK, int MESH>
hipError_t adjoint_mixed_impl(jgm_context* ctx, int kind, int nbm, const double* rm,
                              size_t rm_ld, double* grad, size_t grad_ld, bool with_bg) {
  constexpr int NC = Ncomp<MASK>::value;
  hipStream_t st = ctx->stream;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const FarGeom g = far_geom(ctx);
  NearBoxes& nb = ctx->near;
  hipError_t e;
  if (N > 0) {
    const size_t SB = (size_t)S * nbm;
    if ((e = ctx->coef.ensure(sizeof(double) * SB * NC)) != hipSuccess) return e;
    if ((e = ctx->coef32.ensure(sizeof(float) * SB * NC)) != hipSuccess) return e;
    if ((e = ctx->coef_mono.ensure(sizeof(float) * SB * 8)) != hipSuccess) return e;
    for (int b = 0; b < nbm; ++b)
      hipLaunchKernelGGL((k_mixed_coef<MASK>), dim3(blocks_for(S)), dim3(kBlock), 0, st,
                         rm + b * rm_ld, S, kind, field_dir(ctx),
                         ctx->coef.as<double>() + (size_t)b * S * NC,
                         ctx->coef32.as<float>() + (size_t)b * S * NC,
                         ctx->coef_mono.as<float>() + (size_t)b * S * 8);
    if ((e = ensure_shape_table<MESH>(ctx)) != hipSuccess) return e;
    // the adjoint walks 4x4x(4 kQ) tiles, kQ cells per lane
    const long ntiles8 = (long)g.ntx * g.nty * ((g.nzf + kQ * kTile - 1) / (kQ * kTile));
    if ((e = ensure_far_plan<MESH>(ctx)) != hipSuccess) return e;
    const int nchunks = nb.adj_chunks;
    if ((e = ctx->hpart.ensure(sizeof(double) * (size_t)nbm * nchunks * N)) != hipSuccess)
      return e;
    if ((e = ctx->hsum.ensure(sizeof(double) * (size_t)N)) != hipSuccess) return e;
    {
      ProfScope ps(ctx, far_name(MASK, true));
      launch_far_adjoint<MASK, MESH>(nbm, dim3(blocks_for(ntiles8 * 64), nchunks), st,
                                     ctx->d_sx.as<const double>(), ctx->d_sy.as<const double>(),
                                     ctx->d_sz.as<const double>(), ctx->coef32.as<const float>(),
                                     ctx->coef_mono.as<const float>(), nb.box.as<const int>(),
                                     second_radius_q2(ctx), S, nchunks,
                                     nb.adj_ranges.as<const int>(), g,
                                     nb.adj_table.as<const float>(), uniform_const(ctx), ntiles8,
                                     ctx->hpart.as<double>());
    }
    if (nb.ncols > 0) {
      if ((e = ctx->near_h.ensure(sizeof(double) * std::max<size_t>(1, nb.total_near_corners))) !=
          hipSuccess)
        return e;
      if ((e = ctx->near_v.ensure(sizeof(double) * std::max<size_t>(1, nb.total_near_cells))) !=
          hipSuccess)
        return e;
    }
    for (int b = 0; b < nbm; ++b) {
      hipLaunchKernelGGL(k_fill_zero, dim3(blocks_for(N)), dim3(kBlock), 0, st,
                         ctx->hsum.as<double>(), (size_t)N);
      if (nb.ncols > 0) {
        ProfScope ps(ctx, "near_adjoint_fp64");
        hipLaunchKernelGGL((k_near_adjoint_site<MASK>), dim3(S), dim3(kBlock), 0, st,
                           ctx->d_sx.as<double>(), ctx->d_sy.as<double>(),
                           ctx->d_sz.as<double>(), nb.box.as<const int>(),
                           ctx->coef.as<const double>() + (size_t)b * S * NC,
                           ctx->d_Xc.as<double>(), ctx->d_Yc.as<double>(),
                           ctx->d_Zc.as<double>(), nb.corner_off.as<const long>(),
                           nb.cell_off.as<const long>(), ctx->near_h.as<double>(),
                           ctx->near_v.as<double>());
        hipLaunchKernelGGL(k_near_adjoint_gather, dim3((unsigned)nb.ncols), dim3(64), 0, st,
                           nb.col_ptr.as<const int>(), nb.col_sites.as<const int>(),
                           nb.col_ids.as<const int>(), nb.col_kmax.as<const int>(),
                           nb.box.as<const int>(), nb.cell_off.as<const long>(),
                           ctx->near_v.as<const double>(), g.ny, g.nz, ctx->hsum.as<double>());
      }
      hipLaunchKernelGGL(k_adjoint_mixed_finish, dim3(blocks_for(N)), dim3(kBlock), 0, st,
                         ctx->hpart.as<const double>() + (size_t)b * nchunks * N, nchunks,
                         ctx->hsum.as<const double>(), N,
                         nb.far_subset ? nb.listed.as<const int>() : nullptr, g.nz,
                         grad + b * grad_ld);
    }
  }
  if (with_bg && !ctx->bgthick.empty()) {
    for (int b = 0; b < nbm; ++b) {
      e = launch_background_gradient(ctx, kind, rm + b * rm_ld, grad + b * grad_ld + N);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

// The z-column mask serves far-layer subsets only, which never take the cube paths.
template <int MASK>
hipError_t forward_dispatch(jgm_context* ctx, int kind, int nbm, const double* models,
                            size_t model_ld, double* out, size_t out_ld, double* raw_out,
                            size_t raw_ld, double* sums = nullptr) {
  switch (ctx->near.mesh_class) {
    case kMeshCubes:
      if constexpr (MASK == kMaskZcol)
        return hipErrorNotSupported;
      else
        return forward_mixed_impl<MASK, kMeshCubes>(ctx, kind, nbm, models, model_ld, out,
                                                    out_ld, raw_out, raw_ld, sums);
    case kMeshUniform:
      return forward_mixed_impl<MASK, kMeshUniform>(ctx, kind, nbm, models, model_ld, out, out_ld,
                                                    raw_out, raw_ld, sums);
    default:
      return forward_mixed_impl<MASK, kMeshGeneral>(ctx, kind, nbm, models, model_ld, out, out_ld,
                                                    raw_out, raw_ld, sums);
  }
}

template <int MASK>
hipError_t adjoint_dispatch(jgm_context* ctx, int kind, int nbm, const double* rm, size_t rm_ld,
                            double* grad, size_t grad_ld, bool with_bg = true) {
  switch (ctx->near.mesh_class) {
    case kMeshCubes:
      if constexpr (MASK == kMaskZcol)
        return hipErrorNotSupported;
      else
        return adjoint_mixed_impl<MASK, kMeshCubes>(ctx, kind, nbm, rm, rm_ld, grad, grad_ld,
                                                    with_bg);
    case kMeshUniform:
      return adjoint_mixed_impl<MASK, kMeshUniform>(ctx, kind, nbm, rm, rm_ld, grad, grad_ld,
                                                    with_bg);
    default:
      return adjoint_mixed_impl<MASK, kMeshGeneral>(ctx, kind, nbm, rm, rm_ld, grad, grad_ld,
                                                    with_bg);
  }
}

// 1e9 mu0 / (4 pi G): magnetization in A/m to nT through the tensor terms (which carry G)
double magnetization_unit() {
  const double mu0 = 4e-7 * kPi;  // OMPMagnetizationImp.cpp:35
  return 1.0e9 * mu0 / (4 * kPi * kGravConst);
}

// Magnetization kind in mixed precision: the tensor far field of the weight sets Mx, My,
// Mz (one group of three) shares each pair's multipole terms; near field per set.
hipError_t magnetization_forward_mixed(jgm_context* ctx, const double* model, double* out,
                                       double* raw_out) {
  const int S = (int)ctx->S;
  const size_t N = ncells(ctx);
  hipError_t e = ctx->mag_part.ensure(sizeof(double) * 3 * 6 * std::max(1, S));
  if (e != hipSuccess) return e;
  double* sums = ctx->mag_part.as<double>();
  e = forward_dispatch<kMaskTensor>(ctx, JGM_KIND_TENSOR, 3, model, N, nullptr, 0, nullptr, 0,
                                    sums);
  if (e != hipSuccess) return e;
  const double tf_bx = std::cos(ctx->tf_inc) * std::cos(ctx->tf_dec);
  const double tf_by = std::cos(ctx->tf_inc) * std::sin(ctx->tf_dec);
  const double tf_bz = std::sin(ctx->tf_inc);
  hipLaunchKernelGGL(k_mag3_mixed_finish, dim3(blocks_for(S)), dim3(kBlock), 0, ctx->stream,
                     ctx->mag_part.as<const double>(), S, magnetization_unit(), ctx->transform,
                     tf_bx, tf_by, tf_bz, out, raw_out);
  return hipGetLastError();
}

hipError_t magnetization_adjoint_mixed(jgm_context* ctx, const double* rm, double* grad) {
  const int S = (int)ctx->S;
  const size_t N = ncells(ctx);
  hipError_t e = ctx->mag_part.ensure(sizeof(double) * 27 * std::max(1, S));
  if (e != hipSuccess) return e;
  double* rmt = ctx->mag_part.as<double>();
  hipLaunchKernelGGL(k_mag3_tensor_misfit, dim3(blocks_for(27L * S)), dim3(kBlock), 0,
                     ctx->stream, rm, S, magnetization_unit(), rmt);
  return adjoint_dispatch<kMaskTensor>(ctx, JGM_KIND_TENSOR, 3, rmt, 9 * (size_t)S, grad, N,
                                       false);
}

}  // namespace

hipError_t launch_fill_zero(jgm_context* ctx, double* p, size_t n) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     ctx->stream, p, n);
  return hipGetLastError();
}

// Host-side planning of the near/far split (geometry only): per-site index boxes,
// the per-column site lists and per-site buffer offsets of the adjoint near pass,
// and the far-field geometry.
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
  // reach of layer k: q0 half-diagonals of the largest cell shape of the layer.  A site's
  // box takes the layers whose centre is within their reach of it, and along x and y the
  // largest reach of those layers, so every cell outside the box is at least q0 of its own
  // half-diagonals away (along z if its layer was not taken, else along x or y)
  std::vector<double> reach(nz);
  for (int k = 0; k < nz; ++k)
    reach[k] = ctx->mixed_q0 * 0.5 * std::sqrt(mx * mx + my * my + ctx->zs[k] * ctx->zs[k]);
  nb.radius = ctx->mixed_q0 * 0.5 * std::sqrt(mx * mx + my * my + mz * mz);
  // the far field's layers: the listed ones (increasing mesh layers), else every layer
  const bool subset = !ctx->far_layers.empty();
  std::vector<int> fl;
  if (subset) {
    fl = ctx->far_layers;
    for (size_t f = 0; f < fl.size(); ++f)
      if (fl[f] < 0 || fl[f] >= nz || (f > 0 && fl[f] <= fl[f - 1])) return hipErrorInvalidValue;
  } else {
    for (int k = 0; k < nz; ++k) fl.push_back(k);
  }
  const int nzf = (int)fl.size();
  std::vector<double> zcen_f(nzf), zs_f(nzf);
  for (int f = 0; f < nzf; ++f) {
    zcen_f[f] = zcen[fl[f]];
    zs_f[f] = ctx->zs[fl[f]];
  }
  // the cell shapes the far field sees; a listed subset never takes the cube paths, which
  // run over every layer
  bool uniform = nx > 0 && ny > 0 && nzf > 0;
  for (double v : ctx->xs) uniform = uniform && v == ctx->xs[0];
  for (double v : ctx->ys) uniform = uniform && v == ctx->ys[0];
  for (double v : zs_f) uniform = uniform && v == zs_f[0];
  const bool cubes = !subset && uniform && ctx->xs[0] == ctx->ys[0] && ctx->xs[0] == zs_f[0];
  nb.mesh_class = cubes ? kMeshCubes : uniform ? kMeshUniform : kMeshGeneral;
  nb.far_subset = subset;
  nb.nzf = nzf;
  nb.far_dz = nzf > 0 ? zs_f[0] : 0.0;
  // monotone centres are required by the binary searches
  bool monotone = true;
  for (int i = 1; i < nx; ++i) monotone = monotone && xcen[i] > xcen[i - 1];
  for (int j = 1; j < ny; ++j) monotone = monotone && ycen[j] > ycen[j - 1];
  for (int k = 1; k < nz; ++k) monotone = monotone && zcen[k] > zcen[k - 1];
  if (!monotone) return hipErrorInvalidValue;
  std::vector<int> box(std::max<size_t>(1, S) * 6);
  std::vector<int> count((size_t)nx * ny + 1, 0);
  std::vector<long> corner_off(std::max<size_t>(1, S), 0), cell_off(std::max<size_t>(1, S), 0);
  size_t total_cells = 0, total_corners = 0;
  for (size_t s = 0; s < S; ++s) {
    int lo[3], hi[3];
    double Rh = 0.0;
    lo[2] = nz;
    hi[2] = -1;
    for (int k = 0; k < nz; ++k)
      if (std::fabs(ctx->hsz[s] - zcen[k]) < reach[k]) {
        lo[2] = std::min(lo[2], k);
        hi[2] = k;
        Rh = std::max(Rh, reach[k]);
      }
    axis_range(xcen, ctx->hsx[s], Rh, &lo[0], &hi[0]);
    axis_range(ycen, ctx->hsy[s], Rh, &lo[1], &hi[1]);
    bool empty = false;
    for (int a = 0; a < 3; ++a) empty = empty || lo[a] > hi[a];
    corner_off[s] = (long)total_corners;
    cell_off[s] = (long)total_cells;
    if (empty) {
      lo[0] = lo[1] = lo[2] = 1;
      hi[0] = hi[1] = hi[2] = 0;
    } else {
      const size_t di = hi[0] - lo[0] + 1, dj = hi[1] - lo[1] + 1, dk = hi[2] - lo[2] + 1;
      total_cells += di * dj * dk;
      total_corners += (di + 1) * (dj + 1) * (dk + 1);
      for (int i = lo[0]; i <= hi[0]; ++i)
        for (int j = lo[1]; j <= hi[1]; ++j) count[(size_t)i * ny + j]++;
    }
    for (int a = 0; a < 3; ++a) {
      box[s * 6 + 2 * a] = lo[a];
      box[s * 6 + 2 * a + 1] = hi[a];
    }
  }
  nb.total_near_cells = total_cells;
  nb.total_near_corners = total_corners;
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
  hipStream_t st = ctx->stream;
  hipError_t e;
  if ((e = upload_vec(nb.box, box, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_ptr, col_ptr, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_sites, col_sites, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_ids, col_ids, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.col_kmax, col_kmax, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.corner_off, corner_off, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.cell_off, cell_off, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.xcen, xcen, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.ycen, ycen, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.zcen, zcen_f, st)) != hipSuccess) return e;
  if (subset) {
    std::vector<int> listed(nz, 0);
    for (int k : fl) listed[k] = 1;
    if ((e = upload_vec(nb.kmap, fl, st)) != hipSuccess ||
        (e = upload_vec(nb.listed, listed, st)) != hipSuccess)
      return e;
  }
  auto run_offsets = [](const std::vector<double>& cen) {
    std::vector<float> off(cen.size());
    for (size_t i = 0; i < cen.size(); ++i) off[i] = (float)(cen[i] - cen[i / kTile * kTile]);
    return off;
  };
  const std::vector<float> xoff = run_offsets(xcen), yoff = run_offsets(ycen),
                           zoff = run_offsets(zcen_f);
  // squared largest half-size of each run of `len` cells along an axis
  auto run_h2 = [](const std::vector<double>& sz, int len) {
    std::vector<float> h((sz.size() + len - 1) / len + 1, 0.0f);
    for (size_t i = 0; i < sz.size(); ++i)
      h[i / len] = std::max(h[i / len], (float)(0.25 * sz[i] * sz[i]));
    return h;
  };
  if ((e = upload_vec(nb.hx2, run_h2(ctx->xs, kTile), st)) != hipSuccess ||
      (e = upload_vec(nb.hy2, run_h2(ctx->ys, kTile), st)) != hipSuccess ||
      (e = upload_vec(nb.hz2, run_h2(zs_f, kTile), st)) != hipSuccess ||
      (e = upload_vec(nb.hz2_16, run_h2(zs_f, kQ * kTile), st)) != hipSuccess)
    return e;
  if ((e = upload_vec(nb.xoff, xoff, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.yoff, yoff, st)) != hipSuccess) return e;
  if ((e = upload_vec(nb.zoff, zoff, st)) != hipSuccess) return e;
  const std::vector<int> order = morton_order(ctx);
  if ((e = upload_vec(nb.site_order, order, st)) != hipSuccess) return e;
  nb.plan_valid = false;  // the far adjoint's partition is made again on its next call
  // sites with two coordinates on corner planes of the mesh lie on a mesh line: there the
  // reference's guarded logarithms (log(max(u + r, eps)), BasicGravElements.cpp:53-59)
  // give every cell with an edge on that line a term that the far-field expansion does not
  // have, so these sites take the FP64 closed form over the whole mesh
  nb.line_sites.clear();
  for (size_t s = 0; s < S; ++s) {
    const int on = (int)std::binary_search(ctx->xc.begin(), ctx->xc.end(), ctx->hsx[s]) +
                   (int)std::binary_search(ctx->yc.begin(), ctx->yc.end(), ctx->hsy[s]) +
                   (int)std::binary_search(ctx->zc.begin(), ctx->zc.end(), ctx->hsz[s]);
    if (on >= 2) nb.line_sites.push_back((int)s);
  }
  nb.nline = (int)nb.line_sites.size();
  if (nb.nline > 0) {
    std::vector<double> xyz(3 * (size_t)nb.nline);
    for (int f = 0; f < nb.nline; ++f) {
      xyz[f] = ctx->hsx[nb.line_sites[f]];
      xyz[nb.nline + f] = ctx->hsy[nb.line_sites[f]];
      xyz[2 * (size_t)nb.nline + f] = ctx->hsz[nb.line_sites[f]];
    }
    if ((e = upload_vec(nb.line_idx, nb.line_sites, st)) != hipSuccess) return e;
    if ((e = upload_vec(nb.line_xyz, xyz, st)) != hipSuccess) return e;
  }
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  nb.adj_table_valid = false;
  nb.valid = true;
  return hipSuccess;
}

namespace {
hipError_t route_line_sites_forward(jgm_context* ctx, int kind, size_t nmodels,
                                    const double* models, size_t model_ld, double* out,
                                    size_t out_ld, double* raw_out, size_t raw_ld) {
  const NearBoxes& nb = ctx->near;
  if (nb.nline == 0 || ncells(ctx) == 0) return hipSuccess;
  const double* xyz = nb.line_xyz.as<const double>();
  const int F = nb.nline;
  for (size_t b = 0; b < nmodels; ++b) {
    hipError_t e = launch_forward_fp64_sites(ctx, kind, models + b * model_ld, F, xyz, xyz + F,
                                             xyz + 2 * F, nb.line_idx.as<const int>(),
                                             out + b * out_ld,
                                             raw_out ? raw_out + b * raw_ld : nullptr);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
}  // namespace

hipError_t launch_zcol_sums_mixed(jgm_context* ctx, const double* rho, double* sums) {
  if (!ctx->near.valid) {
    hipError_t e = build_near_boxes(ctx);
    if (e != hipSuccess) return e;
  }
  return forward_dispatch<kMaskZcol>(ctx, JGM_KIND_TENSOR, 1, rho, ncells(ctx), nullptr, 0,
                                     nullptr, 0, sums);
}

hipError_t launch_zcol_adjoint_mixed(jgm_context* ctx, const double* rmt, double* grad) {
  if (!ctx->near.valid) {
    hipError_t e = build_near_boxes(ctx);
    if (e != hipSuccess) return e;
  }
  return adjoint_dispatch<kMaskZcol>(ctx, JGM_KIND_TENSOR, 1, rmt, 9 * ctx->S, grad,
                                     ncells(ctx), false);
}

hipError_t launch_forward_mixed(jgm_context* ctx, int kind, const double* model, double* out,
                                double* raw_out) {
  return launch_forward_mixed_batch(ctx, kind, 1, model, 0, out, 0, raw_out, 0);
}

hipError_t launch_forward_mixed_batch(jgm_context* ctx, int kind, size_t nmodels,
                                      const double* models, size_t model_ld, double* out,
                                      size_t out_ld, double* raw_out, size_t raw_ld) {
  if (kind == JGM_KIND_MAGNETIC_GRADIENT) {  // through the slab mesh's child context
    for (size_t b = 0; b < nmodels; ++b) {
      hipError_t e = launch_forward_maggrad_mixed(ctx, models + b * model_ld, out + b * out_ld,
                                                  raw_out ? raw_out + b * raw_ld : nullptr);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  if (!ctx->near.valid) {
    hipError_t e = build_near_boxes(ctx);
    if (e != hipSuccess) return e;
  }
  if (kind == JGM_KIND_MAGNETIZATION) {  // three weight sets per model
    for (size_t b = 0; b < nmodels; ++b) {
      hipError_t e = magnetization_forward_mixed(ctx, models + b * model_ld, out + b * out_ld,
                                                 raw_out ? raw_out + b * raw_ld : nullptr);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  for (size_t b0 = 0; b0 < nmodels;) {
    const size_t left = nmodels - b0;
    const int nbm = left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const double* m = models + b0 * model_ld;
    double* o = out + b0 * out_ld;
    double* r = raw_out ? raw_out + b0 * raw_ld : nullptr;
    hipError_t e;
    switch (mask_of(kind)) {
      case kMaskScalar:
        e = forward_dispatch<kMaskScalar>(ctx, kind, nbm, m, model_ld, o, out_ld, r, raw_ld);
        break;
      case kMaskAll:
        e = forward_dispatch<kMaskAll>(ctx, kind, nbm, m, model_ld, o, out_ld, r, raw_ld);
        break;
      default:
        e = forward_dispatch<kMaskTensor>(ctx, kind, nbm, m, model_ld, o, out_ld, r, raw_ld);
        break;
    }
    if (e != hipSuccess) return e;
    b0 += nbm;
  }
  // sites on mesh lines: the FP64 closed form in place of the mixed result
  return route_line_sites_forward(ctx, kind, nmodels, models, model_ld, out, out_ld, raw_out,
                                  raw_ld);
}

hipError_t launch_adjoint_mixed(jgm_context* ctx, int kind, const double* /*model*/,
                                const double* raw_misfit, double* grad) {
  return launch_adjoint_mixed_batch(ctx, kind, 1, raw_misfit, 0, grad, 0);
}

hipError_t launch_adjoint_mixed_batch(jgm_context* ctx, int kind, size_t nmodels,
                                      const double* raw_misfits, size_t rm_ld, double* grads,
                                      size_t grad_ld) {
  if (kind == JGM_KIND_MAGNETIC_GRADIENT) {
    for (size_t b = 0; b < nmodels; ++b) {
      hipError_t e =
          launch_adjoint_maggrad_mixed(ctx, raw_misfits + b * rm_ld, grads + b * grad_ld);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  if (!ctx->near.valid) {
    hipError_t e = build_near_boxes(ctx);
    if (e != hipSuccess) return e;
  }
  if (kind == JGM_KIND_MAGNETIZATION) {
    for (size_t b = 0; b < nmodels; ++b) {
      hipError_t e =
          magnetization_adjoint_mixed(ctx, raw_misfits + b * rm_ld, grads + b * grad_ld);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  // sites on mesh lines: their misfits are taken out of the mixed adjoint and their
  // cell gradients added from the FP64 closed form below
  const NearBoxes& nbx = ctx->near;
  const bool route = nbx.nline > 0 && ncells(ctx) > 0;
  const double* all_misfits = raw_misfits;
  const size_t all_ld = rm_ld;
  if (route) {
    const size_t kraw = raw_per_site(kind), n = ctx->S * kraw;
    hipError_t e = ctx->rt_rm0.ensure(sizeof(double) * n * nmodels);
    if (e != hipSuccess) return e;
    double* rm0 = ctx->rt_rm0.as<double>();
    for (size_t b = 0; b < nmodels; ++b) {
      if ((e = hipMemcpyAsync(rm0 + b * n, raw_misfits + b * rm_ld, sizeof(double) * n,
                              hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess ||
          (e = launch_zero_site_rows(ctx, rm0 + b * n, nbx.line_idx.as<const int>(), nbx.nline,
                                     (int)kraw)) != hipSuccess)
        return e;
    }
    raw_misfits = rm0;
    rm_ld = n;
  }
  for (size_t b0 = 0; b0 < nmodels;) {
    const size_t left = nmodels - b0;
    const int nbm = left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const double* r = raw_misfits + b0 * rm_ld;
    double* gr = grads + b0 * grad_ld;
    hipError_t e;
    switch (mask_of(kind)) {
      case kMaskScalar:
        e = adjoint_dispatch<kMaskScalar>(ctx, kind, nbm, r, rm_ld, gr, grad_ld);
        break;
      case kMaskAll: e = adjoint_dispatch<kMaskAll>(ctx, kind, nbm, r, rm_ld, gr, grad_ld); break;
      default: e = adjoint_dispatch<kMaskTensor>(ctx, kind, nbm, r, rm_ld, gr, grad_ld); break;
    }
    if (e != hipSuccess) return e;
    b0 += nbm;
  }
  if (route) {
    const double* xyz = nbx.line_xyz.as<const double>();
    const int F = nbx.nline;
    const long N = (long)ncells(ctx);
    for (size_t b = 0; b < nmodels; ++b) {
      double* g = grads + b * grad_ld;
      const double* rm = all_misfits + b * all_ld;
      hipError_t e;
      // the background gradient of every site, theirs included
      if (!ctx->bgthick.empty() &&
          (e = launch_background_gradient(ctx, kind, rm, g + N)) != hipSuccess)
        return e;
      if ((e = launch_adjoint_fp64_sites_add(ctx, kind, F, xyz, xyz + F, xyz + 2 * F,
                                             nbx.line_idx.as<const int>(), rm, g)) != hipSuccess)
        return e;
    }
  }
  return hipSuccess;
}

}  // namespace jgm
