// Per-site epilogues shared by the FP64 and mixed-precision paths: background,
// assembly of the raw data layout, magnetic projection, data transforms, and the
// per-site coefficients of the adjoint.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "context.hpp"
#include "prism_device.hpp"

namespace jgm {

struct FieldDir {
  double bx, by, bz, scale;  // direction cosines and F / (4 pi G)
};

__host__ inline FieldDir field_dir(const jgm_context* ctx) {
  FieldDir f;
  f.bx = std::cos(ctx->inclination) * std::cos(ctx->declination);
  f.by = std::cos(ctx->inclination) * std::sin(ctx->declination);
  f.bz = std::sin(ctx->inclination);
  f.scale = ctx->fieldstrength * (1.0 / (4 * kPi * kGravConst));
  return f;
}

// u[] holds gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz of the gridded part (physical values).
// Adds the background layers and assembles the raw data of the kind into raw[] (gz;
// the 9 tensor values row-major, TensorOMPGravityImp.cpp:255-264; or
// B = (T b) F/(4 pi G), OMPMagneticSusceptibilityImp.cpp:164-170); returns its length.
__device__ inline int assemble_raw(double* u, int s, const double* __restrict__ bgT,
                                   const double* __restrict__ bgvals, int L, int kind,
                                   FieldDir fd, double* raw) {
  for (int j = 0; j < L; ++j) {
    const double* b = bgT + ((long)s * L + j) * 7;
    const double v = bgvals[j];
    for (int m = 0; m < 7; ++m) u[m] += b[m] * v;
  }
  if (kind == JGM_KIND_SCALAR) {
    raw[0] = u[0];
    return 1;
  }
  if (kind == JGM_KIND_TENSOR || kind == JGM_KIND_SCALAR_TENSOR) {
    int o = 0;
    if (kind == JGM_KIND_SCALAR_TENSOR) raw[o++] = u[0];
    raw[o + 0] = u[1]; raw[o + 1] = u[4]; raw[o + 2] = u[5];
    raw[o + 3] = u[4]; raw[o + 4] = u[2]; raw[o + 5] = u[6];
    raw[o + 6] = u[5]; raw[o + 7] = u[6]; raw[o + 8] = u[3];
    return o + 9;
  }
  raw[0] = (u[1] * fd.bx + u[4] * fd.by + u[5] * fd.bz) * fd.scale;
  raw[1] = (u[4] * fd.bx + u[2] * fd.by + u[6] * fd.bz) * fd.scale;
  raw[2] = (u[5] * fd.bx + u[6] * fd.by + u[3] * fd.bz) * fd.scale;
  return 3;
}

// Stores the raw data of site s (if raw_out) and the transformed data into out.
__device__ inline void store_site(const double* raw, int kraw, int s, int tf, double tfbx,
                                  double tfby, double tfbz, double* __restrict__ out,
                                  double* __restrict__ raw_out) {
  if (raw_out) {
    for (int m = 0; m < kraw; ++m) raw_out[(long)s * kraw + m] = raw[m];
  }
  if (tf == JGM_TF_FTG_INVARIANT) {
    const double* d = raw;  // GravityTransforms.h:36-40
    out[s] = d[0] * d[4] + d[4] * d[8] + d[0] * d[8] - d[3] * d[1] - d[7] * d[5] - d[2] * d[6];
  } else if (tf == JGM_TF_TOTAL_FIELD) {
    out[s] = sqrt(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2]);
  } else if (tf == JGM_TF_TOTAL_FIELD_ANOMALY) {
    out[s] = (tfbx * raw[0] + tfby * raw[1] + tfbz * raw[2]);
  } else {
    for (int m = 0; m < kraw; ++m) out[(long)s * kraw + m] = raw[m];
  }
}

// Background, raw assembly, transform and store for one site.
__device__ inline void finish_site(double* u, int s, const double* __restrict__ bgT,
                                   const double* __restrict__ bgvals, int L, int kind,
                                   FieldDir fd, int tf, double tfbx, double tfby, double tfbz,
                                   double* __restrict__ out, double* __restrict__ raw_out) {
  double raw[10];
  const int kraw = assemble_raw(u, s, bgT, bgvals, L, kind, fd, raw);
  store_site(raw, kraw, s, tf, tfbx, tfby, tfbz, out, raw_out);
}

// Coefficients of the 7 physical field components (gz, Uxx, Uyy, Uzz, Uxy, Uxz, Uyz)
// in sum_k rraw[s][k] * raw_k: the adjoint of the assembly above.
__device__ inline void site_coefficients(const double* __restrict__ rm, int s, int kind,
                                         FieldDir fd, double* c) {
  for (int m = 0; m < 7; ++m) c[m] = 0.0;
  if (kind == JGM_KIND_SCALAR) {
    c[0] = rm[s];
  } else if (kind == JGM_KIND_TENSOR || kind == JGM_KIND_SCALAR_TENSOR) {
    const int k = kind == JGM_KIND_SCALAR_TENSOR ? 10 : 9;
    const double* r = rm + (long)s * k;
    int o = 0;
    if (kind == JGM_KIND_SCALAR_TENSOR) c[0] = r[o++];
    c[1] = r[o + 0];
    c[2] = r[o + 4];
    c[3] = r[o + 8];
    c[4] = r[o + 1] + r[o + 3];
    c[5] = r[o + 2] + r[o + 6];
    c[6] = r[o + 5] + r[o + 7];
  } else {
    const double* r = rm + (long)s * 3;
    const double f = fd.scale;
    c[1] = r[0] * fd.bx * f;
    c[2] = r[1] * fd.by * f;
    c[3] = r[2] * fd.bz * f;
    c[4] = (r[0] * fd.by + r[1] * fd.bx) * f;
    c[5] = (r[0] * fd.bz + r[2] * fd.bx) * f;
    c[6] = (r[1] * fd.bz + r[2] * fd.by) * f;
  }
}

// Select the components of MASK from the 7 (kernel component order).
template <int MASK>
__host__ __device__ inline void select_components(const double* c7, double* out) {
  int k = 0;
  if (MASK & kZcol) { out[k++] = c7[3]; out[k++] = c7[5]; out[k++] = c7[6]; }
  if (MASK & kGz) out[k++] = c7[0];
  if (MASK & kDiag) { out[k++] = c7[1]; out[k++] = c7[2]; out[k++] = c7[3]; }
  if (MASK & kOff) { out[k++] = c7[4]; out[k++] = c7[5]; out[k++] = c7[6]; }
}

// Inverse of select_components: spread MASK components into the 7-slot layout.
template <int MASK>
__host__ __device__ inline void spread_components(const double* in, double* u7) {
  int k = 0;
  for (int m = 0; m < 7; ++m) u7[m] = 0.0;
  if (MASK & kZcol) { u7[3] = in[k++]; u7[5] = in[k++]; u7[6] = in[k++]; }
  if (MASK & kGz) u7[0] = in[k++];
  if (MASK & kDiag) { u7[1] = in[k++]; u7[2] = in[k++]; u7[3] = in[k++]; }
  if (MASK & kOff) { u7[4] = in[k++]; u7[5] = in[k++]; u7[6] = in[k++]; }
}

}  // namespace jgm
