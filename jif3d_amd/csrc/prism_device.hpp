// Device-side prism mathematics for gfx950.
//
// The closed form of jif3D's box terms is an 8-corner alternating sum of a corner
// function F(x, y, z) of the site-to-corner offset (BasicGravElements.cpp:26-59,
// :82-244).  Every corner of the mesh is shared by up to eight cells with the same
// sign pattern sigma(a,b,c) = (-1)^(a+b+c+1) for the corner at offset (a,b,c) of a
// cell.  So for one site
//
//   sum_c rho_c J[s,c] = sign_m * G * sum_corners F_m(s, corner) * w(corner),
//   w(I,J,K) = sum_{a,b,c in {0,1}} sigma(a,b,c) rho(I-a, J-b, K-c),
//
// and for the adjoint g_c = 2 sum_s r_s J[s,c] = 2 sum_{corners of c} sigma H(corner)
// with H(corner) = sum_s sum_m coef[s][m] F_m(s, corner).  Each corner function is
// evaluated once per (site, corner) pair: (nx+1)(ny+1)(nz+1) evaluations per site
// instead of 8 nx ny nz.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace jgm {

constexpr double kGravConst = 6.67428e-11;  // jif3D/Gravity/BasicGravElements.h:22
constexpr double kPi = 3.14159265358979323846;

// Component masks of the corner function.  Component order inside kernels:
// gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz.  kZcol alone (never combined with the others) is the
// tensor's z column, Uzz, Uxz, Uyz: the terms the vertical magnetic gradient needs.
enum : int { kGz = 1, kDiag = 2, kOff = 4, kZcol = 8 };
constexpr int kMaskScalar = kGz;
constexpr int kMaskTensor = kDiag | kOff;
constexpr int kMaskAll = kGz | kDiag | kOff;
constexpr int kMaskZcol = kZcol;

template <int MASK>
struct Ncomp {
  static constexpr int value = ((MASK & kGz) ? 1 : 0) + ((MASK & kDiag) ? 3 : 0) +
                               ((MASK & kOff) ? 3 : 0) + ((MASK & kZcol) ? 3 : 0);
};

// Natural logarithm of a finite a >= DBL_EPSILON (the guarded log arguments of the
// corner function), written for this use: a = 2^e m with m in [1/sqrt2, sqrt2),
// log m = 2 atanh(f) = 2 f + 2 f f^2 q(f^2), f = (m - 1)/(m + 1), |f| <= 0.1716, q the
// degree-7 minimax polynomial of jif3d_amd/csrc/corner_poly.py (q ~ 1/3 + f^2/5 + ...;
// relative error 3e-18, below 1e-19 in the logarithm).  f comes from the refined hardware reciprocal
// and is corrected by its exact residual (m + 1 split into a sum of two doubles), and the
// result is assembled in double-double, so the error stays near half an ulp; about 45
// VALU operations against ~95 for the library log, which also handles zero, negative,
// subnormal and non-finite arguments that cannot occur here.
__device__ __forceinline__ double log_pos_f64(double a) {
  double m = __builtin_amdgcn_frexp_mant(a);  // [0.5, 1)
  int e = __builtin_amdgcn_frexp_exp(a);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double num = m - 1.0;  // exact
  const double den = m + 1.0;
  const double bb0 = den - m;
  const double dlo = (m - (den - bb0)) + (1.0 - bb0);  // den + dlo = m + 1 exactly
  // hardware reciprocal estimate (relative error below 2^-24.4, measured on gfx950) and one
  // Newton step (2.2e-15); f's remaining error is carried by flo, formed from the exact residual
  double rc = __builtin_amdgcn_rcp(den);
  rc = fma(rc, fma(-den, rc, 1.0), rc);
  const double f = num * rc;
  const double flo = (fma(-f, den, num) - f * dlo) * rc;  // f + flo = num / (m + 1)
  const double f2 = f * f;
  double q = 0.0654368673815417;
  q = fma(q, f2, 0.06634346805939063);
  q = fma(q, f2, 0.07693121530786205);
  q = fma(q, f2, 0.09090897784329731);
  q = fma(q, f2, 0.11111111196258087);
  q = fma(q, f2, 0.14285714285399548);
  q = fma(q, f2, 0.20000000000000445);
  q = fma(q, f2, 0.3333333333333333);
  q *= f2;  // ~ f^2/3 + f^4/5 + ...
  const double ln2_hi = 6.93147180369123816490e-01;  // 32 significant bits: e ln2_hi exact
  const double ln2_lo = 1.90821492927058770002e-10;
  const double de = (double)e;
  const double hi = de * ln2_hi;
  const double t = 2.0 * f;
  const double sum = hi + t;
  const double bb = sum - hi;
  const double err = (hi - (sum - bb)) + (t - bb);
  const double lo = fma(t, q, fma(de, ln2_lo, 2.0 * flo)) + err;
  return sum + lo;
}

// Arctangent of p / q in the quadrant of (q, p), atan2(p, q), for p and q of magnitudes in
// [1e-280, 1e280] (the corner terms' products of site-corner offsets in [1e-140, 1e140];
// corner_terms_f64 sends other offsets to the library atan2), written for this use: the octant reduction
// u = min / max of |p|, |q| and v = u, or (u - 1) / (u + 1) when u > tan(pi/8), from one
// division (hardware reciprocal, one Newton step, the quotient's residual), atan(v) = v + v^3 g(v^2) with the
// degree-10 minimax polynomial g of jif3d_amd/csrc/corner_poly.py (relative error 1.3e-18), and
// the octant and quadrant offset m pi/4 added in two parts (m hi exact for m <= 4).  Within
// 2 ulp of the exact value over 2e5 random arguments of every octant; ~45 VALU operations
// against ~70 for the library atan2, which also handles zero, infinite and NaN arguments.
__device__ __forceinline__ double atan2_f64(double p, double q) {
  const double ap = fabs(p), aq = fabs(q);
  const bool swap = ap > aq;
  const double num = swap ? aq : ap, den = swap ? ap : aq;
  const bool big = num > 0.41421356237309503 * den;  // tan(pi/8)
  const double n2 = big ? num - den : num, d2 = big ? num + den : den;
  double r = __builtin_amdgcn_rcp(d2);
  r = fma(r, fma(-d2, r, 1.0), r);  // one Newton step: 2.2e-15 relative (log_pos_f64)
  double v = n2 * r;
  v = fma(r, fma(-d2, v, n2), v);  // the quotient corrected by its residual
  const double x = v * v;
  double g = -0.01790504451210959;
  g = fma(g, x, 0.03806214458319016);
  g = fma(g, x, -0.05039190615403013);
  g = fma(g, x, 0.05847859316652799);
  g = fma(g, x, -0.06663099209930878);
  g = fma(g, x, 0.0769205971799347);
  g = fma(g, x, -0.09090897725284844);
  g = fma(g, x, 0.11111110782156222);
  g = fma(g, x, -0.14285714280166587);
  g = fma(g, x, 0.19999999999953247);
  g = fma(g, x, -0.333333333333332);
  const double at = fma(v * x, g, v);
  int m = big ? 1 : 0;
  m = swap ? 2 - m : m;
  m = q < 0.0 ? 4 - m : m;
  const double s = (swap != (q < 0.0)) ? -1.0 : 1.0;
  const double dm = (double)m;
  const double res = fma(dm, 0.7853981633974483, fma(dm, 3.061616997868383e-17, s * at));
  return p < 0.0 ? -res : res;
}

// Corner function in FP64.  x, y, z = site - corner.
//  gz  : x log(y+r) + y log(x+r) + z atan2(z r, x y), r floored at DBL_EPSILON (:26-37)
//  Uxx : atan2(y z, x r); Uyy: atan2(x z, y r); Uzz: atan2(x y, z r)         (:43-47, :223-229)
//  Uxy : log(max(z + r, eps)); Uxz: log(max(y + r, eps)); Uyz: log(max(x + r, eps))
//                                                                            (:53-59, :231-241)
// With the gravity term alone, the logs keep the unguarded form of :32 so that a
// site on a mesh line reproduces the reference bit for bit in its special cases;
// with the tensor terms present gz reuses the guarded logs of the off-diagonal
// terms (they coincide unless x = z = 0 and y < 0, where the reference's gz term
// is 0 * log(0) = NaN).
//
// Two identities of the atan2 terms save two of the four atan2 evaluations when the
// diagonal terms are present and no coordinate is zero:
//  atan2(yz, xr) + atan2(xz, yr) + atan2(xy, zr) = (pi/2) * {1, 1, -3, 5}[number of
//    negative coordinates]  (the solid-angle identity of a rectangular corner, with the
//    quadrant offsets of atan2), and
//  atan2(zr, xy) + atan2(xy, zr) = -3pi/2 if xy < 0 and z < 0, pi/2 otherwise.
// The term of the coordinate smallest in magnitude (the largest angle) is the one
// derived, so the subtraction does not cancel.
template <int MASK>
__device__ __forceinline__ void diag_terms_f64(double x, double y, double z, double r,
                                               double* A, double* B, double* C) {
  const double ax = fabs(x), ay = fabs(y), az = fabs(z);
  const int which = (ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2);  // derived term
  const double p1 = which == 0 ? x * z : y * z, q1 = which == 0 ? y * r : x * r;
  const double p2 = which == 2 ? x * z : x * y, q2 = which == 2 ? y * r : z * r;
  const double t1 = atan2_f64(p1, q1), t2 = atan2_f64(p2, q2);
  const int neg = (x < 0.0) + (y < 0.0) + (z < 0.0);
  const double K = neg <= 1 ? 0.5 * kPi : (neg == 2 ? -1.5 * kPi : 2.5 * kPi);
  const double t3 = (K - t1) - t2;
  *A = which == 0 ? t3 : t1;
  *B = which == 0 ? t1 : (which == 1 ? t3 : t2);
  *C = which == 2 ? t3 : t2;
}

template <int MASK>
__device__ __forceinline__ void corner_terms_f64(double x, double y, double z, double* t) {
  const double r = sqrt(x * x + y * y + z * z);
  if constexpr (MASK == kMaskZcol) {
    // Uzz = atan2(x y, z r), Uxz = log(max(y + r, eps)), Uyz = log(max(x + r, eps)); the
    // atan2 term evaluated directly (the tensor mask may derive it from its two others)
    t[0] = (fmin(fmin(fabs(x), fabs(y)), fabs(z)) >= 1e-140 && r >= DBL_EPSILON && r <= 1e140)
               ? atan2_f64(x * y, z * r)
               : atan2(x * y, z * r);
    t[1] = log_pos_f64(fmax(y + r, DBL_EPSILON));
    t[2] = log_pos_f64(fmax(x + r, DBL_EPSILON));
    return;
  }
  int k = 0;
  if constexpr ((MASK & kOff) != 0) {
    const double lz = log_pos_f64(fmax(z + r, DBL_EPSILON));
    const double ly = log_pos_f64(fmax(y + r, DBL_EPSILON));
    const double lx = log_pos_f64(fmax(x + r, DBL_EPSILON));
    if constexpr ((MASK & kDiag) != 0) {
      double A, B, C, G;
      // no zero coordinate, and every product of two offsets within atan2_f64's range
      if (fmin(fmin(fabs(x), fabs(y)), fabs(z)) >= 1e-140 && r >= DBL_EPSILON && r <= 1e140) {
        diag_terms_f64<MASK>(x, y, z, r, &A, &B, &C);
        G = ((x * y < 0.0 && z < 0.0) ? -1.5 * kPi : 0.5 * kPi) - C;
      } else {
        A = atan2(y * z, x * r);
        B = atan2(x * z, y * r);
        C = atan2(x * y, z * r);
        G = atan2(z * fmax(r, DBL_EPSILON), x * y);
      }
      if constexpr ((MASK & kGz) != 0) t[k++] = x * ly + y * lx + z * G;
      t[k++] = A;
      t[k++] = B;
      t[k++] = C;
    } else if constexpr ((MASK & kGz) != 0) {
      const double rg = fmax(r, DBL_EPSILON);
      double v = x * ly + y * lx;
      v += z * atan2(z * rg, x * y);
      t[k++] = v;
    }
    t[k++] = lz;
    t[k++] = ly;
    t[k++] = lx;
  } else {
    if constexpr ((MASK & kGz) != 0) {
      const double rg = fmax(r, DBL_EPSILON);
      double v = x * log(y + rg) + y * log(x + rg);
      v += z * atan2(z * rg, x * y);
      t[k++] = v;
    }
    if constexpr ((MASK & kDiag) != 0) {
      t[k++] = atan2(y * z, x * r);
      t[k++] = atan2(x * z, y * r);
      t[k++] = atan2(x * y, z * r);
    }
  }
}

// Sign and scale applied to each accumulated corner sum, in the kernel's
// component order: gz and the off-diagonal (log) terms are returned as -G * sum
// (BasicGravElements.cpp:108, :201), the diagonal (atan2) terms as +G * sum (:151).
template <int MASK>
__host__ __device__ inline void comp_scales(double* s) {
  int k = 0;
  if (MASK & kZcol) { s[k++] = kGravConst; s[k++] = -kGravConst; s[k++] = -kGravConst; }
  if (MASK & kGz) s[k++] = -kGravConst;
  if (MASK & kDiag) { s[k++] = kGravConst; s[k++] = kGravConst; s[k++] = kGravConst; }
  if (MASK & kOff) { s[k++] = -kGravConst; s[k++] = -kGravConst; s[k++] = -kGravConst; }
}

// Full per-cell box term (8 corners) for the background boxes and the
// per-cell sensitivity rows.  Returns sign-applied, G-scaled values in the
// kernel's component order, like CalcGravBoxTerm / CalcTensorBoxTerm.
template <int MASK>
__device__ inline void box_terms_f64(double mx, double my, double mz, double ux, double uy,
                                     double uz, double sx, double sy, double sz, double* out) {
  constexpr int NC = Ncomp<MASK>::value;
  double acc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = 0.0;
  const double xs[2] = {mx - ux, mx - (ux + sx)};
  const double ys[2] = {my - uy, my - (uy + sy)};
  const double zs[2] = {mz - uz, mz - (uz + sz)};
  // corners in the reference's summation order (BasicGravElements.cpp:88-106)
  constexpr int order[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                               {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int a = order[n][0], b = order[n][1], c = order[n][2];
    double t[NC];
    corner_terms_f64<MASK>(xs[a], ys[b], zs[c], t);
    if (n == 0) {
#pragma unroll
      for (int m = 0; m < NC; ++m) acc[m] = -t[m];
    } else if ((a + b + c) & 1) {
#pragma unroll
      for (int m = 0; m < NC; ++m) acc[m] += t[m];
    } else {
#pragma unroll
      for (int m = 0; m < NC; ++m) acc[m] -= t[m];
    }
  }
  double sc[NC];
  comp_scales<MASK>(sc);
#pragma unroll
  for (int m = 0; m < NC; ++m) out[m] = acc[m] * sc[m];
}

}  // namespace jgm
