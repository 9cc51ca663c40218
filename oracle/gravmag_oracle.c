/*
 * gravmag_oracle.c -- CPU oracle for the jif3D prism gravity / FTG / magnetic path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke() entry
 * point and bench.py's CPU baseline leg may load it; the product library
 * (jif3d_amd) never links, loads or calls it.
 *
 * It is a from-scratch FP64 restatement of the reference algorithm, with the same
 * operand order, the same epsilon guards and the same loop structure (serial loop
 * over sites, OpenMP reduction over the flattened, z-fastest cell index), plus an
 * 80-bit long double variant of the forward problem used as a truth oracle.
 * Reference citations are relative to the mmoorkamp/jif3D tree.
 *
 *   a1  CalcGravTerm                 jif3D/Gravity/BasicGravElements.cpp:26-37
 *   a2  CalcGravBoxTerm              jif3D/Gravity/BasicGravElements.cpp:82-109
 *   a3  CalcFTGDiagonalTerm/CalcUxxTerm   BasicGravElements.cpp:43-47, :125-152
 *   a4  CalcFTGOffDiagonalTerm/CalcUxyTerm BasicGravElements.cpp:53-59, :168-202
 *   a5  CalcTensorBoxTerm            jif3D/Gravity/BasicGravElements.cpp:216-244
 *   a6  CalcInfSheetTerm / CalcUzzInfSheetTerm  jif3D/Gravity/BasicGravElements.h:51-75
 *   a7  CalcScalarBackground / CalcTensorBackground  jif3D/Gravity/GravityBackground.cpp:24-137
 *   a8  ScalarOMPGravityImp::CalcGridded      jif3D/Gravity/ScalarOMPGravityImp.cpp:138-180
 *   a9  TensorOMPGravityImp::CalcGridded      jif3D/Gravity/TensorOMPGravityImp.cpp:125-182
 *   a10 OMPMagneticSusceptibilityImp          jif3D/Magnetics/OMPMagneticSusceptibilityImp.cpp:31-194
 *   a11 OMPMagneticGradImp::CalcGridded       jif3D/Magnetics/OMPMagneticGradImp.cpp:13-106
 *   a12 OMPMagnetizationImp::CalcGridded      jif3D/Magnetics/OMPMagnetizationImp.cpp:25-117
 *   a13 ThreeDGravMagImplementation::Calculate    jif3D/GravMag/ThreeDGravMagImplementation.h:116-168
 *   a14 ThreeDGravMagImplementation::LQDerivative jif3D/GravMag/ThreeDGravMagImplementation.h:177-231
 *   a15 FullSensitivity J.m and 2 J^T r      jif3D/GravMag/FullSensitivityGravMagCalculator.h:137-221
 *   a17 FTGInvariant / TotalField / TotalFieldAnomaly
 *                                    jif3D/Gravity/GravityTransforms.h:25-100,
 *                                    jif3D/Magnetics/MagneticTransforms.h:19-156
 *
 * How each part is pinned (tests/test_oracle_*.py): the reference's known-answer
 * constants (Li & Chouteau top-of-cube value, far-point value, symmetry,
 * translation invariance; jif3D/Gravity/test_ScalarCalculation.cpp:32-64), the
 * discretisation-independence test (:67-120), the analytic slab and layered
 * background (:123-244), tensor trace/symmetry/Poisson (test_TensorCalculation.cpp:
 * 32-57, :187-227), tensor = finite difference of gz (:60-106), the SimPEG golden
 * vectors for the magnetic imp (jif3D/magx.simpeg etc., 0.2 %) and the analytic
 * 2-D magnetic block (test_MagneticCalculation.cpp:28-108), LQDerivative = 2 J^T r
 * (test_ScalarCalculation.cpp:275-306), finite differences of the misfit, the
 * transform test values (jif3D/Gravity/test_transforms.cpp), a brute-force
 * Gauss-Legendre quadrature of the Newton kernel on small boxes, and the
 * long double restatement.  a11 and a12 have no fixtures of their own (the
 * reference's *_vector.simpeg files are absent); they are pinned through the
 * SimPEG-pinned susceptibility imp: the vertical gradient equals 4 pi 1e-7 times the
 * difference of its Bz at z and z + 1 m, and the magnetization imp with
 * M = chi F b_hat 1e-9/mu0 (one axis at a time, then oblique) equals it.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORC_G 6.67428e-11 /* jif3D/Gravity/BasicGravElements.h:22 */
#define ORC_PI 3.14159265358979323846264338327950288

enum { ORC_SCALAR = 1, ORC_TENSOR = 2, ORC_MAGNETIC = 3, ORC_MAGGRAD = 4, ORC_MAGNETIZATION = 5 };
enum { ORC_TF_NONE = 0, ORC_TF_FTG_INVARIANT = 1, ORC_TF_TOTAL_FIELD = 2, ORC_TF_TOTAL_FIELD_ANOMALY = 3 };

/* Planted mistakes for the tests that show the transform-chain pins can fail: built
 * only with -DORC_MUTANT=n (oracle/Makefile target liboracle_mutant<n>.so); 0, the
 * default, is the faithful restatement.
 *   1  TotalField Jacobian B/|B|^2 instead of B/|B|
 *   2  FTGInvariant Jacobian drops U22 from dI/dU00
 *   3  TotalFieldAnomaly Jacobian with cos(dec) and sin(dec) exchanged
 *   4  transformed LQDerivative contracts the Jacobian with the wrong J row
 *   5  TotalField value drops the Bz term
 *   6  transformed LQDerivative without the factor 2 */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* jif3D::sign, jif3D/Global/NumUtil.h:27-33 */
static inline int orc_sign(double n) { return n >= 0 ? 1 : -1; }

/* ------------------------------------------------------------------------- */
/* Box terms.  A macro instantiates them for double and long double so that the */
/* truth oracle follows exactly the same formulas.                             */
/* ------------------------------------------------------------------------- */
#define ORC_DEFINE_BOX(T, SFX, LOGF, ATAN2F, SQRTF, EPS)                                     \
  /* a1: x log(y+r) + y log(x+r) + z atan2(z r, x y); r floored at epsilon (:26-37) */        \
  static inline T grav_term_##SFX(T x, T y, T z) {                                            \
    T r = SQRTF(x * x + y * y + z * z);                                                       \
    if (r < (T)(EPS)) r = (T)(EPS);                                                           \
    T rvalue = x * LOGF(y + r) + y * LOGF(x + r);                                             \
    rvalue += z * ATAN2F(z * r, x * y);                                                       \
    return rvalue;                                                                            \
  }                                                                                           \
  /* a3: atan2(a b, c r) (:43-47) */                                                          \
  static inline T ftg_diag_##SFX(T a, T b, T c) {                                             \
    return ATAN2F(a * b, c * SQRTF(a * a + b * b + c * c));                                   \
  }                                                                                           \
  /* a4: log(max(value + r, eps)) (:53-59) */                                                 \
  static inline T ftg_offdiag_##SFX(T value, T x, T y, T z) {                                 \
    T v = value + SQRTF(x * x + y * y + z * z);                                               \
    if (v < (T)(EPS)) v = (T)(EPS);                                                           \
    return LOGF(v);                                                                           \
  }                                                                                           \
  /* a2 (:82-109) */                                                                          \
  static T grav_box_##SFX(T mx, T my, T mz, T ux, T uy, T uz, T sx, T sy, T sz) {             \
    T rv = -grav_term_##SFX(mx - ux, my - uy, mz - uz);                                       \
    rv += grav_term_##SFX(mx - (ux + sx), my - uy, mz - uz);                                  \
    rv += grav_term_##SFX(mx - ux, my - (uy + sy), mz - uz);                                  \
    rv += grav_term_##SFX(mx - ux, my - uy, mz - (uz + sz));                                  \
    rv -= grav_term_##SFX(mx - (ux + sx), my - (uy + sy), mz - uz);                           \
    rv -= grav_term_##SFX(mx - (ux + sx), my - uy, mz - (uz + sz));                           \
    rv -= grav_term_##SFX(mx - ux, my - (uy + sy), mz - (uz + sz));                           \
    rv += grav_term_##SFX(mx - (ux + sx), my - (uy + sy), mz - (uz + sz));                    \
    return -rv * (T)ORC_G;                                                                    \
  }                                                                                           \
  /* a3 (:125-152) */                                                                         \
  static T uxx_##SFX(T mx, T my, T mz, T ux, T uy, T uz, T sx, T sy, T sz) {                  \
    T rv = -ftg_diag_##SFX(my - uy, mz - uz, mx - ux);                                        \
    rv += ftg_diag_##SFX(my - uy, mz - uz, mx - (ux + sx));                                   \
    rv += ftg_diag_##SFX(my - (uy + sy), mz - uz, mx - ux);                                   \
    rv += ftg_diag_##SFX(my - uy, mz - (uz + sz), mx - ux);                                   \
    rv -= ftg_diag_##SFX(my - (uy + sy), mz - uz, mx - (ux + sx));                            \
    rv -= ftg_diag_##SFX(my - uy, mz - (uz + sz), mx - (ux + sx));                            \
    rv -= ftg_diag_##SFX(my - (uy + sy), mz - (uz + sz), mx - ux);                            \
    rv += ftg_diag_##SFX(my - (uy + sy), mz - (uz + sz), mx - (ux + sx));                     \
    return rv * (T)ORC_G;                                                                     \
  }                                                                                           \
  /* a4 (:168-202) */                                                                         \
  static T uxy_##SFX(T mx, T my, T mz, T ux, T uy, T uz, T sx, T sy, T sz) {                  \
    T rv = -ftg_offdiag_##SFX(mz - uz, my - uy, mz - uz, mx - ux);                            \
    rv += ftg_offdiag_##SFX(mz - uz, my - uy, mz - uz, mx - (ux + sx));                       \
    rv += ftg_offdiag_##SFX(mz - uz, my - (uy + sy), mz - uz, mx - ux);                       \
    rv += ftg_offdiag_##SFX(mz - (uz + sz), my - uy, mz - (uz + sz), mx - ux);                \
    rv -= ftg_offdiag_##SFX(mz - uz, my - (uy + sy), mz - uz, mx - (ux + sx));                \
    rv -= ftg_offdiag_##SFX(mz - (uz + sz), my - uy, mz - (uz + sz), mx - (ux + sx));         \
    rv -= ftg_offdiag_##SFX(mz - (uz + sz), my - (uy + sy), mz - (uz + sz), mx - ux);         \
    rv += ftg_offdiag_##SFX(mz - (uz + sz), my - (uy + sy), mz - (uz + sz), mx - (ux + sx));  \
    return -rv * (T)ORC_G;                                                                    \
  }                                                                                           \
  /* a5 (:216-244); out is the row-major 3x3 matrix */                                       \
  static void tensor_box_##SFX(T mx, T my, T mz, T ux, T uy, T uz, T sx, T sy, T sz,          \
                               T out[9]) {                                                    \
    out[0] = uxx_##SFX(mx, my, mz, ux, uy, uz, sx, sy, sz);                                   \
    out[4] = uxx_##SFX(my, mx, mz, uy, ux, uz, sy, sx, sz);                                   \
    out[8] = uxx_##SFX(mz, mx, my, uz, ux, uy, sz, sx, sy);                                   \
    out[1] = uxy_##SFX(mx, my, mz, ux, uy, uz, sx, sy, sz);                                   \
    out[3] = out[1];                                                                          \
    out[2] = uxy_##SFX(mx, mz, my, ux, uz, uy, sx, sz, sy);                                   \
    out[6] = out[2];                                                                          \
    out[5] = uxy_##SFX(mz, my, mx, uz, uy, ux, sz, sy, sx);                                   \
    out[7] = out[5];                                                                          \
  }

ORC_DEFINE_BOX(double, d, log, atan2, sqrt, DBL_EPSILON)
ORC_DEFINE_BOX(long double, ld, logl, atan2l, sqrtl, DBL_EPSILON)

/* a6: jif3D/Gravity/BasicGravElements.h:51-63 */
static double inf_sheet_d(double measz, double top, double bottom) {
  const double thick = bottom - top;
  if ((measz <= top) || (measz >= bottom))
    return 2.0 * ORC_PI * ORC_G * thick * orc_sign(top - measz);
  return 2.0 * ORC_PI * ORC_G * (thick - 2.0 * (measz - top));
}
/* a6: jif3D/Gravity/BasicGravElements.h:65-75 */
static double uzz_inf_sheet_d(double measz, double top, double bottom) {
  if ((measz <= top) || (measz >= bottom)) return 0.0;
  return 4.0 * ORC_PI * ORC_G;
}
static long double inf_sheet_ld(long double measz, long double top, long double bottom) {
  const long double thick = bottom - top;
  const long double pi = 3.14159265358979323846264338327950288L;
  if ((measz <= top) || (measz >= bottom))
    return 2.0L * pi * (long double)ORC_G * thick * (top - measz >= 0 ? 1 : -1);
  return 2.0L * pi * (long double)ORC_G * (thick - 2.0L * (measz - top));
}
static long double uzz_inf_sheet_ld(long double measz, long double top, long double bottom) {
  if ((measz <= top) || (measz >= bottom)) return 0.0L;
  return 4.0L * 3.14159265358979323846264338327950288L * (long double)ORC_G;
}

/* ------------------------------------------------------------------------- */
/* Exported single-prism functions (used by the known-answer tests).           */
/* ------------------------------------------------------------------------- */
double orc_grav_box_term(double mx, double my, double mz, double ux, double uy, double uz,
                         double sx, double sy, double sz) {
  return grav_box_d(mx, my, mz, ux, uy, uz, sx, sy, sz);
}
void orc_tensor_box_term(double mx, double my, double mz, double ux, double uy, double uz,
                         double sx, double sy, double sz, double *out9) {
  tensor_box_d(mx, my, mz, ux, uy, uz, sx, sy, sz, out9);
}
double orc_grav_box_term_ld(double mx, double my, double mz, double ux, double uy, double uz,
                            double sx, double sy, double sz) {
  return (double)grav_box_ld(mx, my, mz, ux, uy, uz, sx, sy, sz);
}
void orc_tensor_box_term_ld(double mx, double my, double mz, double ux, double uy, double uz,
                            double sx, double sy, double sz, double *out9) {
  long double t[9];
  tensor_box_ld(mx, my, mz, ux, uy, uz, sx, sy, sz, t);
  for (int i = 0; i < 9; ++i) out9[i] = (double)t[i];
}
double orc_inf_sheet_term(double measz, double top, double bottom) {
  return inf_sheet_d(measz, top, bottom);
}
double orc_uzz_inf_sheet_term(double measz, double top, double bottom) {
  return uzz_inf_sheet_d(measz, top, bottom);
}
int orc_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Mesh and survey description shared by the whole-model entry points.         */
/* ------------------------------------------------------------------------- */
typedef struct {
  size_t nx, ny, nz;
  const double *xc, *yc, *zc; /* GridX/Y/ZCoordinates, n+1 entries, jif3D/ModelBase/ThreeDModelBase.cpp:96-113 */
  const double *xs, *ys, *zs; /* X/Y/ZCellSizes, n entries */
  size_t nbg;
  const double *bgthick;      /* background layer thicknesses */
} orc_mesh;

typedef struct {
  double inclination, declination, fieldstrength; /* radians, radians, nT */
} orc_field;

/* rows of raw data per site: ndatapermeas of the jif3D imps (ScalarOMPGravityImp.h:31, TensorOMPGravityImp.h:39, OMPMagneticSusceptibilityImp.h:41) */
static size_t raw_per_site(int kind) {
  switch (kind) {
  case ORC_SCALAR: return 1;
  case ORC_TENSOR: return 9;
  case ORC_MAGNETIC: return 3;
  case ORC_MAGGRAD: return 1;
  case ORC_MAGNETIZATION: return 3;
  default: return 0;
  }
}
/* model parameters: grid (+ 3x grid for magnetization) plus background layers */
static size_t n_params(int kind, const orc_mesh *m) {
  size_t n = m->nx * m->ny * m->nz;
  if (kind == ORC_MAGNETIZATION) return 3 * n; /* ThreeDMagnetizationModel, no background */
  if (kind == ORC_MAGGRAD) return n + m->nbg;
  return n + m->nbg;
}

/* jif3D/ModelBase/ThreeDModelBase.h:162-168 */
static inline void offset_to_index(const orc_mesh *m, int offset, int *xi, int *yi, int *zi) {
  *zi = offset % (int)m->nz;
  *xi = (offset - *zi) / (int)m->nz;
  *yi = *xi % (int)m->ny;
  *xi = (*xi - *yi) / (int)m->ny;
}

/* ------------------------------------------------------------------------- */
/* Gridded part for one site.  sens (nullable) has raw_per_site rows of ld      */
/* columns.  Returns the raw data in out[].                                     */
/* ------------------------------------------------------------------------- */
static void gridded(int kind, const orc_mesh *m, const orc_field *f, const double *model,
                    double sx, double sy, double sz, double *out, double *sens, size_t ld) {
  const int nmod = (int)(m->nx * m->ny * m->nz);
  if (kind == ORC_SCALAR) {
    /* a8: jif3D/Gravity/ScalarOMPGravityImp.cpp:153-180 */
    double returnvalue = 0.0;
#pragma omp parallel for reduction(+ : returnvalue)
    for (int offset = 0; offset < nmod; ++offset) {
      int xi, yi, zi;
      offset_to_index(m, offset, &xi, &yi, &zi);
      double cur = grav_box_d(sx, sy, sz, m->xc[xi], m->yc[yi], m->zc[zi], m->xs[xi], m->ys[yi],
                              m->zs[zi]);
      returnvalue += cur * model[offset];
      if (sens) sens[offset] = cur;
    }
    out[0] = returnvalue;
  } else if (kind == ORC_TENSOR) {
    /* a9: jif3D/Gravity/TensorOMPGravityImp.cpp:145-182 */
    double U0 = 0, U1 = 0, U2 = 0, U3 = 0, U4 = 0, U5 = 0, U6 = 0, U7 = 0, U8 = 0;
#pragma omp parallel for reduction(+ : U0, U1, U2, U3, U4, U5, U6, U7, U8)
    for (int offset = 0; offset < nmod; ++offset) {
      int xi, yi, zi;
      double c[9];
      offset_to_index(m, offset, &xi, &yi, &zi);
      tensor_box_d(sx, sy, sz, m->xc[xi], m->yc[yi], m->zc[zi], m->xs[xi], m->ys[yi], m->zs[zi],
                   c);
      const double d = model[offset];
      U0 += c[0] * d; U1 += c[1] * d; U2 += c[2] * d;
      U3 += c[3] * d; U4 += c[4] * d; U5 += c[5] * d;
      U6 += c[6] * d; U7 += c[7] * d; U8 += c[8] * d;
      if (sens)
        for (int i = 0; i < 9; ++i) sens[(size_t)i * ld + offset] = c[i];
    }
    out[0] = U0; out[1] = U1; out[2] = U2; out[3] = U3; out[4] = U4;
    out[5] = U5; out[6] = U6; out[7] = U7; out[8] = U8;
  } else if (kind == ORC_MAGNETIC) {
    /* a10: jif3D/Magnetics/OMPMagneticSusceptibilityImp.cpp:113-194 */
    const double BxComp = cos(f->inclination) * cos(f->declination);
    const double ByComp = cos(f->inclination) * sin(f->declination);
    const double BzComp = sin(f->inclination);
    const double fs = f->fieldstrength;
    double Bx = 0, By = 0, Bz = 0;
#pragma omp parallel for reduction(+ : Bx, By, Bz)
    for (int offset = 0; offset < nmod; ++offset) {
      int xi, yi, zi;
      double c[9];
      offset_to_index(m, offset, &xi, &yi, &zi);
      tensor_box_d(sx, sy, sz, m->xc[xi], m->yc[yi], m->zc[zi], m->xs[xi], m->ys[yi], m->zs[zi],
                   c);
      const double chi = model[offset];
      const double factor = 1.0 / (4 * ORC_PI * ORC_G);
      const double BxS = (c[0] * BxComp + c[1] * ByComp + c[2] * BzComp) * fs * factor;
      const double ByS = (c[3] * BxComp + c[4] * ByComp + c[5] * BzComp) * fs * factor;
      const double BzS = (c[6] * BxComp + c[7] * ByComp + c[8] * BzComp) * fs * factor;
      Bx += BxS * chi; By += ByS * chi; Bz += BzS * chi;
      if (sens) {
        sens[offset] = BxS;
        sens[ld + offset] = ByS;
        sens[2 * ld + offset] = BzS;
      }
    }
    out[0] = Bx; out[1] = By; out[2] = Bz;
  } else if (kind == ORC_MAGGRAD) {
    /* a11: jif3D/Magnetics/OMPMagneticGradImp.cpp:13-106 */
    const double BxComp = cos(f->inclination) * cos(f->declination);
    const double ByComp = cos(f->inclination) * sin(f->declination);
    const double BzComp = sin(f->inclination);
    const double fs = f->fieldstrength;
    double BzBot = 0, BzTop = 0;
#pragma omp parallel for reduction(+ : BzBot, BzTop)
    for (int offset = 0; offset < nmod; ++offset) {
      int xi, yi, zi;
      double cb[9], ct[9];
      offset_to_index(m, offset, &xi, &yi, &zi);
      tensor_box_d(sx, sy, sz, m->xc[xi], m->yc[yi], m->zc[zi], m->xs[xi], m->ys[yi], m->zs[zi],
                   cb);
      const double chi = model[offset];
      const double factor = 1e-7 / ORC_G;
      const double BzSensBot = (cb[6] * BxComp + cb[7] * ByComp + cb[8] * BzComp) * fs * factor;
      BzBot += BzSensBot * chi;
      tensor_box_d(sx, sy, sz + 1, m->xc[xi], m->yc[yi], m->zc[zi], m->xs[xi], m->ys[yi],
                   m->zs[zi], ct);
      const double BzSensTop = (ct[6] * BxComp + ct[7] * ByComp + ct[8] * BzComp) * fs * factor;
      BzTop += BzSensTop * chi;
      if (sens) sens[offset] = BzSensBot - BzSensTop;
    }
    out[0] = BzBot - BzTop;
  } else if (kind == ORC_MAGNETIZATION) {
    /* a12: jif3D/Magnetics/OMPMagnetizationImp.cpp:25-117; model = [Mx; My; Mz] */
    const double mag_mu = 4e-7 * ORC_PI;
    double Bx = 0, By = 0, Bz = 0;
#pragma omp parallel for reduction(+ : Bx, By, Bz)
    for (int offset = 0; offset < nmod; ++offset) {
      int xi, yi, zi;
      double c[9];
      offset_to_index(m, offset, &xi, &yi, &zi);
      const double factor = 1.0e9 * mag_mu / (4 * ORC_PI * ORC_G);
      tensor_box_d(sx, sy, sz, m->xc[xi], m->yc[yi], m->zc[zi], m->xs[xi], m->ys[yi], m->zs[zi],
                   c);
      for (int i = 0; i < 9; ++i) c[i] = factor * c[i];
      const double mx = model[offset], my = model[offset + nmod], mz = model[offset + 2 * nmod];
      Bx += (c[0] * mx + c[1] * my + c[2] * mz);
      By += (c[3] * mx + c[4] * my + c[5] * mz);
      Bz += (c[6] * mx + c[7] * my + c[8] * mz);
      if (sens) {
        for (int r = 0; r < 3; ++r) {
          sens[(size_t)r * ld + offset] = c[3 * r + 0];
          sens[(size_t)r * ld + offset + nmod] = c[3 * r + 1];
          sens[(size_t)r * ld + offset + 2 * nmod] = c[3 * r + 2];
        }
      }
    }
    out[0] = Bx; out[1] = By; out[2] = Bz;
  }
}

/* ------------------------------------------------------------------------- */
/* 1-D background for one site (a7 and the magnetic background of a10).         */
/* out[] receives raw_per_site values; sens (nullable) columns nmod..nmod+nbg.  */
/* ------------------------------------------------------------------------- */
static void background(int kind, const orc_mesh *m, const orc_field *f, const double *bgvals,
                       double sx, double sy, double sz, double xwidth, double ywidth,
                       double zwidth, double *out, double *sens, size_t ld) {
  const size_t nmod = m->nx * m->ny * m->nz;
  const size_t k = raw_per_site(kind);
  for (size_t i = 0; i < k; ++i) out[i] = 0.0;
  if (kind == ORC_MAGGRAD || kind == ORC_MAGNETIZATION) return; /* both return zeros */
  double currtop = m->zc[0];
  double currbottom = currtop;
  const double xo = m->xc[0], yo = m->yc[0];
  for (size_t j = 0; j < m->nbg; ++j) {
    const double currthick = m->bgthick[j];
    currbottom = currtop + currthick;
    if (kind == ORC_SCALAR) {
      /* jif3D/Gravity/GravityBackground.cpp:41-68 */
      double currvalue = inf_sheet_d(sz, currtop, currbottom);
      if (currtop < zwidth && (currbottom <= zwidth))
        currvalue -= grav_box_d(sx, sy, sz, xo, yo, currtop, xwidth, ywidth, currthick);
      if (currtop < zwidth && currbottom > zwidth)
        currvalue -= grav_box_d(sx, sy, sz, xo, yo, currtop, xwidth, ywidth, (zwidth - currtop));
      if (sens) sens[nmod + j] = currvalue;
      out[0] += currvalue * bgvals[j];
    } else {
      /* jif3D/Gravity/GravityBackground.cpp:110-133 and
         jif3D/Magnetics/OMPMagneticSusceptibilityImp.cpp:58-98 */
      double cv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, box[9];
      cv[8] = uzz_inf_sheet_d(sz, currtop, currbottom);
      if (currtop < zwidth && (currbottom <= zwidth)) {
        tensor_box_d(sx, sy, sz, xo, yo, currtop, xwidth, ywidth, currthick, box);
        for (int i = 0; i < 9; ++i) cv[i] -= box[i];
      }
      if (currtop < zwidth && currbottom > zwidth) {
        tensor_box_d(sx, sy, sz, xo, yo, currtop, xwidth, ywidth, (zwidth - currtop), box);
        for (int i = 0; i < 9; ++i) cv[i] -= box[i];
      }
      if (kind == ORC_TENSOR) {
        if (sens)
          for (int i = 0; i < 9; ++i) sens[(size_t)i * ld + nmod + j] = cv[i];
        for (int i = 0; i < 9; ++i) out[i] += cv[i] * bgvals[j];
      } else {
        const double BxComp = cos(f->inclination) * cos(f->declination);
        const double ByComp = cos(f->inclination) * sin(f->declination);
        const double BzComp = sin(f->inclination);
        const double factor = 1.0 / (4 * ORC_PI * ORC_G);
        const double fs = f->fieldstrength;
        const double BxS = (cv[0] * BxComp + cv[1] * ByComp + cv[2] * BzComp) * fs * factor;
        const double ByS = (cv[3] * BxComp + cv[4] * ByComp + cv[5] * BzComp) * fs * factor;
        const double BzS = (cv[6] * BxComp + cv[7] * ByComp + cv[8] * BzComp) * fs * factor;
        if (sens) {
          sens[nmod + j] = BxS;
          sens[ld + nmod + j] = ByS;
          sens[2 * ld + nmod + j] = BzS;
        }
        out[0] += BxS * bgvals[j];
        out[1] += ByS * bgvals[j];
        out[2] += BzS * bgvals[j];
      }
    }
    currtop += currthick;
  }
}

/* ------------------------------------------------------------------------- */
/* Data transforms (a17).                                                      */
/* ------------------------------------------------------------------------- */
static size_t tf_out(int tf, size_t kraw) { return tf == ORC_TF_NONE ? kraw : 1; }

static void tf_apply(int tf, double inc, double dec, const double *in, double *out) {
  if (tf == ORC_TF_FTG_INVARIANT) {
    /* jif3D/Gravity/GravityTransforms.h:36-40 */
    out[0] = in[0] * in[4] + in[4] * in[8] + in[0] * in[8] - in[3] * in[1] - in[7] * in[5] -
             in[2] * in[6];
  } else if (tf == ORC_TF_TOTAL_FIELD) {
    /* jif3D/Magnetics/MagneticTransforms.h:27-30 */
    out[0] = ORC_MUTANT == 5 ? sqrt(in[0] * in[0] + in[1] * in[1])
                             : sqrt(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
  } else if (tf == ORC_TF_TOTAL_FIELD_ANOMALY) {
    /* jif3D/Magnetics/MagneticTransforms.h:121-128, constructor :143-149 */
    const double bx = cos(inc) * cos(dec), by = cos(inc) * sin(dec), bz = sin(inc);
    out[0] = (bx * in[0] + by * in[1] + bz * in[2]);
  }
}
/* 1 x kraw Jacobian of the transform at in[] */
static void tf_derivative(int tf, double inc, double dec, const double *in, double *d) {
  if (tf == ORC_TF_FTG_INVARIANT) {
    /* jif3D/Gravity/GravityTransforms.h:80-91 */
    d[0] = ORC_MUTANT == 2 ? in[4] : in[4] + in[8]; d[1] = -in[3]; d[2] = -in[6];
    d[3] = -in[1]; d[4] = in[0] + in[8]; d[5] = -in[7];
    d[6] = -in[2]; d[7] = -in[5]; d[8] = in[0] + in[4];
  } else if (tf == ORC_TF_TOTAL_FIELD) {
    /* jif3D/Magnetics/MagneticTransforms.h:66-72 */
    const double t2 = in[0] * in[0] + in[1] * in[1] + in[2] * in[2];
    const double t = ORC_MUTANT == 1 ? t2 : sqrt(t2);
    d[0] = in[0] / t; d[1] = in[1] / t; d[2] = in[2] / t;
  } else if (tf == ORC_TF_TOTAL_FIELD_ANOMALY) {
    /* jif3D/Magnetics/MagneticTransforms.h:131-141 */
    if (ORC_MUTANT == 3) {
      d[0] = cos(inc) * sin(dec); d[1] = cos(inc) * cos(dec); d[2] = sin(inc);
    } else {
      d[0] = cos(inc) * cos(dec); d[1] = cos(inc) * sin(dec); d[2] = sin(inc);
    }
  }
}

static double accumulate(const double *v, size_t n) {
  double s = 0.0; /* std::accumulate(..., 0.0) */
  for (size_t i = 0; i < n; ++i) s += v[i];
  return s;
}

/* ------------------------------------------------------------------------- */
/* a13: forward for all sites.  out has nsites * kout entries.  J (nullable)   */
/* receives the raw sensitivity matrix (nsites*kraw rows, n_params columns),   */
/* as FullSensitivityGravMagCalculator::HandleSensitivities lays it out        */
/* (jif3D/GravMag/FullSensitivityGravMagCalculator.h:99-114).                  */
/* ------------------------------------------------------------------------- */
int orc_calculate(int kind, size_t nx, size_t ny, size_t nz, const double *xc, const double *yc,
                  const double *zc, const double *xs, const double *ys, const double *zs,
                  size_t nbg, const double *bgthick, const double *model, size_t nsites,
                  const double *sx, const double *sy, const double *sz, double inc, double dec,
                  double fs, int tf, double tf_inc, double tf_dec, double *out, double *J) {
  orc_mesh m = {nx, ny, nz, xc, yc, zc, xs, ys, zs, nbg, bgthick};
  orc_field f = {inc, dec, fs};
  const size_t kraw = raw_per_site(kind);
  if (kraw == 0) return -1;
  const size_t kout = tf_out(tf, kraw);
  const size_t np = n_params(kind, &m);
  const size_t nmod = nx * ny * nz;
  const double xw = accumulate(xs, nx), yw = accumulate(ys, ny), zw = accumulate(zs, nz);
  const double *bgvals = model + nmod;
  double raw[9], bg[9];
  for (size_t i = 0; i < nsites; ++i) {
    double *rows = J ? J + i * kraw * np : NULL;
    gridded(kind, &m, &f, model, sx[i], sy[i], sz[i], raw, rows, np);
    background(kind, &m, &f, bgvals, sx[i], sy[i], sz[i], xw, yw, zw, bg, rows, np);
    for (size_t c = 0; c < kraw; ++c) raw[c] += bg[c];
    if (tf != ORC_TF_NONE)
      tf_apply(tf, tf_inc, tf_dec, raw, out + i * kout);
    else
      for (size_t c = 0; c < kraw; ++c) out[i * kout + c] = raw[c];
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* a14: matrix-free least-squares derivative, one J row block per site.        */
/* ------------------------------------------------------------------------- */
int orc_lqderivative(int kind, size_t nx, size_t ny, size_t nz, const double *xc,
                     const double *yc, const double *zc, const double *xs, const double *ys,
                     const double *zs, size_t nbg, const double *bgthick, const double *model,
                     size_t nsites, const double *sx, const double *sy, const double *sz,
                     double inc, double dec, double fs, int tf, double tf_inc, double tf_dec,
                     const double *misfit, double *grad) {
  orc_mesh m = {nx, ny, nz, xc, yc, zc, xs, ys, zs, nbg, bgthick};
  orc_field f = {inc, dec, fs};
  const size_t kraw = raw_per_site(kind);
  if (kraw == 0) return -1;
  const size_t kout = tf_out(tf, kraw);
  const size_t np = n_params(kind, &m);
  const size_t nmod = nx * ny * nz;
  const double xw = accumulate(xs, nx), yw = accumulate(ys, ny), zw = accumulate(zs, nz);
  const double *bgvals = model + nmod;
  double *rows = (double *)calloc(kraw * np, sizeof(double));
  double *newrow = (double *)malloc(np * sizeof(double));
  if (!rows || !newrow) { free(rows); free(newrow); return -2; }
  for (size_t c = 0; c < np; ++c) grad[c] = 0.0;
  double raw[9], bg[9], d[9];
  for (size_t i = 0; i < nsites; ++i) {
    gridded(kind, &m, &f, model, sx[i], sy[i], sz[i], raw, rows, np);
    background(kind, &m, &f, bgvals, sx[i], sy[i], sz[i], xw, yw, zw, bg, rows, np);
    for (size_t c = 0; c < kraw; ++c) raw[c] += bg[c];
    if (tf != ORC_TF_NONE) {
      /* NewSensitivities = prod(TransDeriv, CurrentSensitivities), :212-219 */
      tf_derivative(tf, tf_inc, tf_dec, raw, d);
      for (size_t c = 0; c < np; ++c) {
        double s = 0.0;
        for (size_t r = 0; r < kraw; ++r)
          s += d[r] * rows[(ORC_MUTANT == 4 ? (r + 1) % kraw : r) * np + c];
        newrow[c] = s;
      }
      const double mf = ORC_MUTANT == 6 ? 0.5 * misfit[i * kout] : misfit[i * kout];
      for (size_t c = 0; c < np; ++c) grad[c] += mf * newrow[c];
    } else {
      /* DerivMod += Misfit(i*k+j) * row j, :224-228 */
      for (size_t j = 0; j < kraw; ++j) {
        const double mf = misfit[i * kout + j];
        const double *row = rows + j * np;
        for (size_t c = 0; c < np; ++c) grad[c] += mf * row[c];
      }
    }
  }
  for (size_t c = 0; c < np; ++c) grad[c] = 2.0 * grad[c];
  free(rows);
  free(newrow);
  return 0;
}

/* a15: d = J m (prec_prod, row by row) */
void orc_j_times(size_t nrows, size_t ncols, const double *J, const double *m, double *out) {
  for (size_t r = 0; r < nrows; ++r) {
    double s = 0.0;
    const double *row = J + r * ncols;
    for (size_t c = 0; c < ncols; ++c) s += row[c] * m[c];
    out[r] = s;
  }
}
/* a15: g = 2 J^T r (prod(trans(J), r), column by column) */
void orc_jt_times2(size_t nrows, size_t ncols, const double *J, const double *r, double *out) {
  for (size_t c = 0; c < ncols; ++c) {
    double s = 0.0;
    for (size_t i = 0; i < nrows; ++i) s += J[i * ncols + c] * r[i];
    out[c] = 2.0 * s;
  }
}

/* ------------------------------------------------------------------------- */
/* Truth oracle: forward in 80-bit long double (scalar, tensor, magnetic), no   */
/* transform.  Same formulas and loop structure; accumulation in long double.   */
/* ------------------------------------------------------------------------- */
int orc_calculate_ld(int kind, size_t nx, size_t ny, size_t nz, const double *xc,
                     const double *yc, const double *zc, const double *xs, const double *ys,
                     const double *zs, size_t nbg, const double *bgthick, const double *model,
                     size_t nsites, const double *sx, const double *sy, const double *sz,
                     double inc, double dec, double fs, double *out) {
  if (kind != ORC_SCALAR && kind != ORC_TENSOR && kind != ORC_MAGNETIC) return -1;
  const int nmod = (int)(nx * ny * nz);
  orc_mesh m = {nx, ny, nz, xc, yc, zc, xs, ys, zs, nbg, bgthick};
  const size_t kraw = raw_per_site(kind);
  long double xw = 0, yw = 0, zw = 0;
  for (size_t i = 0; i < nx; ++i) xw += xs[i];
  for (size_t i = 0; i < ny; ++i) yw += ys[i];
  for (size_t i = 0; i < nz; ++i) zw += zs[i];
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double bxc = cosl(inc) * cosl(dec), byc = cosl(inc) * sinl(dec), bzc = sinl(inc);
  const long double mfac = (long double)fs / (4.0L * pi * (long double)ORC_G);
  for (size_t s = 0; s < nsites; ++s) {
    long double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const long double px = sx[s], py = sy[s], pz = sz[s];
#pragma omp parallel
    {
      long double loc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma omp for
      for (int offset = 0; offset < nmod; ++offset) {
        int xi, yi, zi;
        offset_to_index(&m, offset, &xi, &yi, &zi);
        const long double d = model[offset];
        if (kind == ORC_SCALAR) {
          loc[0] += grav_box_ld(px, py, pz, xc[xi], yc[yi], zc[zi], xs[xi], ys[yi], zs[zi]) * d;
        } else {
          long double c[9];
          tensor_box_ld(px, py, pz, xc[xi], yc[yi], zc[zi], xs[xi], ys[yi], zs[zi], c);
          if (kind == ORC_TENSOR) {
            for (int i = 0; i < 9; ++i) loc[i] += c[i] * d;
          } else {
            loc[0] += (c[0] * bxc + c[1] * byc + c[2] * bzc) * mfac * d;
            loc[1] += (c[3] * bxc + c[4] * byc + c[5] * bzc) * mfac * d;
            loc[2] += (c[6] * bxc + c[7] * byc + c[8] * bzc) * mfac * d;
          }
        }
      }
#pragma omp critical
      for (int i = 0; i < 9; ++i) acc[i] += loc[i];
    }
    /* background in long double */
    long double currtop = zc[0];
    for (size_t j = 0; j < nbg; ++j) {
      const long double th = bgthick[j];
      const long double bot = currtop + th;
      const long double bgv = model[nmod + j];
      if (kind == ORC_SCALAR) {
        long double cv = inf_sheet_ld(pz, currtop, bot);
        if (currtop < zw && bot <= zw) cv -= grav_box_ld(px, py, pz, xc[0], yc[0], currtop, xw, yw, th);
        if (currtop < zw && bot > zw)
          cv -= grav_box_ld(px, py, pz, xc[0], yc[0], currtop, xw, yw, zw - currtop);
        acc[0] += cv * bgv;
      } else {
        long double cv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, box[9];
        cv[8] = uzz_inf_sheet_ld(pz, currtop, bot);
        int have = 0;
        if (currtop < zw && bot <= zw) { tensor_box_ld(px, py, pz, xc[0], yc[0], currtop, xw, yw, th, box); have = 1; }
        if (currtop < zw && bot > zw) { tensor_box_ld(px, py, pz, xc[0], yc[0], currtop, xw, yw, zw - currtop, box); have = 1; }
        if (have) for (int i = 0; i < 9; ++i) cv[i] -= box[i];
        if (kind == ORC_TENSOR) {
          for (int i = 0; i < 9; ++i) acc[i] += cv[i] * bgv;
        } else {
          acc[0] += (cv[0] * bxc + cv[1] * byc + cv[2] * bzc) * mfac * bgv;
          acc[1] += (cv[3] * bxc + cv[4] * byc + cv[5] * bzc) * mfac * bgv;
          acc[2] += (cv[6] * bxc + cv[7] * byc + cv[8] * bzc) * mfac * bgv;
        }
      }
      currtop += th;
    }
    for (size_t c = 0; c < kraw; ++c) out[s * kraw + c] = (double)acc[c];
  }
  return 0;
}

/* Exported transform helpers (a17), used by the transform known-answer tests. */
void orc_transform(int tf, double inc, double dec, const double *in, double *out) {
  tf_apply(tf, inc, dec, in, out);
}
void orc_transform_derivative(int tf, double inc, double dec, const double *in, double *d) {
  tf_derivative(tf, inc, dec, in, d);
}
