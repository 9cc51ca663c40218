// Vertical magnetic gradient (OMPMagneticGradImp.cpp:13-106) in mixed precision.
//
// The reference forms Bz at the site and 1 m below it and differences the two; a far field
// in FP32 cannot do that, the difference being a few per cent of either field for distant
// cells.  The difference is rewritten instead: for a prism P spanning [z1, z2],
//   F_P(s) - F_P(s + 1 m z^) = F_[z1,z2](s) - F_[z1-1,z2-1](s) = F_[z2-1,z2](s) - F_[z1-1,z1](s),
// the fields of two 1 m slabs of P's horizontal extent, one just above its bottom face and
// one just above its top face.  Summed over a column, the slab just above mesh plane K
// (z = Z_K) carries the susceptibility of the cell ending there minus that of the cell
// starting there.  The vertical gradient of the mesh is therefore the Bz of a second mesh
// of nz + 1 slab layers, separated by zero-weight gap layers (Z_K, Z_{K+1} - 1), and that
// mesh goes through the mixed path (FP64 near boxes, FP32 multipole far field) for the
// tensor's z column (Uzz, Uxz, Uyz, the terms Bz needs) of a child context holding its
// geometry, whose far field runs over the slab layers only.  Nothing cancels beyond the slabs' separation, so the far field's FP32 error stays
// at the magnetic kind's level.
//
// The slabs need cells at least 1 m thick; on thinner cells the call runs the FP64 path.
// Sites whose site or 1 m-lower point lies on a mesh line (two coordinates on corner
// planes) take the reference's two-point FP64 form, whose guarded logarithms there
// (BasicGravElements.cpp:53-59) the slab form does not reproduce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "context.hpp"
#include "prism_device.hpp"

namespace jgm {

namespace {

constexpr int kBlock = 256;
inline unsigned blocks_for(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Slab densities: layer kv of the slab mesh holds plane K = plane_of[kv] (or a gap, -1);
// rho_v = chi(K - 1) - chi(K) in each column, chi zero outside the mesh.
__global__ __launch_bounds__(kBlock) void k_slab_weights(const double* __restrict__ chi,
                                                         const int* __restrict__ plane_of,
                                                         long ncol, int nz, int nzv,
                                                         double* __restrict__ rho_v) {
  const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= ncol * nzv) return;
  const int kv = (int)(v % nzv);
  const long col = v / nzv;
  const int K = plane_of[kv];
  double w = 0.0;
  if (K >= 0) {
    const double above = K > 0 ? chi[col * nz + K - 1] : 0.0;
    const double below = K < nz ? chi[col * nz + K] : 0.0;
    w = above - below;
  }
  rho_v[v] = w;
}

// out[s] = F 1e-7 / G (Uxz bx + Uyz by + Uzz bz) of the slab mesh's z-column sums (order
// zz, xz, yz; the sums carry G and the component signs)
__global__ __launch_bounds__(kBlock) void k_slab_finish(const double* __restrict__ sums, int S,
                                                        double bx, double by, double bz,
                                                        double scale, double* __restrict__ out,
                                                        double* __restrict__ raw_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const double* u = sums + (long)s * 3;
  const double v = scale * (u[1] * bx + u[2] * by + u[0] * bz);
  out[s] = v;
  if (raw_out) raw_out[s] = v;
}

// Raw tensor misfit of the slab mesh (order xx, xy, xz, yx, yy, yz, zx, zy, zz): the
// derivative of the data with respect to Uzx, Uzy, Uzz times the misfit
__global__ __launch_bounds__(kBlock) void k_slab_tensor_misfit(const double* __restrict__ rm, int S,
                                                               double bx, double by, double bz,
                                                               double scale,
                                                               double* __restrict__ rmt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9L * S) return;
  const int s = (int)(t / 9), q = (int)(t % 9);
  const double r = rm[s] * scale;
  rmt[t] = q == 6 ? r * bx : q == 7 ? r * by : q == 8 ? r * bz : 0.0;
}

// d/d chi_k of the slab densities: + at the slab of plane k + 1, - at the slab of plane k
__global__ __launch_bounds__(kBlock) void k_slab_gradient(const double* __restrict__ gv,
                                                          const int* __restrict__ layer_of,
                                                          long ncol, int nz, int nzv,
                                                          double* __restrict__ grad) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol * nz) return;
  const int k = (int)(c % nz);
  const long col = c / nz;
  grad[c] = gv[col * nzv + layer_of[k + 1]] - gv[col * nzv + layer_of[k]];
}

double slab_scale(const jgm_context* ctx) { return ctx->fieldstrength * (1e-7 / kGravConst); }

void field_unit(const jgm_context* ctx, double* bx, double* by, double* bz) {
  *bx = std::cos(ctx->inclination) * std::cos(ctx->declination);
  *by = std::cos(ctx->inclination) * std::sin(ctx->declination);
  *bz = std::sin(ctx->inclination);
}

template <class T>
hipError_t upload_vec(DeviceBuffer& b, const std::vector<T>& v, hipStream_t st) {
  hipError_t e = b.ensure(sizeof(T) * std::max<size_t>(1, v.size()));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpyAsync(b.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st);
}

bool on_plane(const std::vector<double>& planes, double p) {
  return std::binary_search(planes.begin(), planes.end(), p);
}

// The slab mesh's child context, its layer maps and the sites on mesh lines, made once
// per mesh and survey (geom_gen); hipErrorNotSupported when a cell is thinner than 1 m.
hipError_t ensure_slab_mesh(jgm_context* ctx) {
  SlabMesh& sm = ctx->slab;
  if (sm.gen == ctx->geom_gen) {
    if (!sm.ok) return hipErrorNotSupported;
  } else {
    sm.gen = ctx->geom_gen;
    sm.ok = false;
    const size_t nz = ctx->nz;
    for (double dz : ctx->zs)
      if (!(dz >= 1.0)) return hipErrorNotSupported;
    // layers: slab of plane 0, then for each cell a gap (when thicker than 1 m) and the
    // slab of the plane below it
    std::vector<double> zlo, zsz;
    std::vector<int> plane_of, layer_of(nz + 1);
    const std::vector<double>& Z = ctx->zc;
    for (size_t K = 0; K <= nz; ++K) {
      if (K > 0) {
        const double gap = (Z[K] - 1.0) - Z[K - 1];
        if (gap > 0.0) {
          zlo.push_back(Z[K - 1]);
          zsz.push_back(gap);
          plane_of.push_back(-1);
        }
      }
      layer_of[K] = (int)zlo.size();
      zlo.push_back(Z[K] - 1.0);
      zsz.push_back(1.0);
      plane_of.push_back((int)K);
    }
    sm.nzv = zlo.size();
    if (!sm.child) {
      jgm_context* c = nullptr;
      if (jgm_create(ctx->device, ctx->stream, &c) != JGM_OK) return hipErrorOutOfMemory;
      sm.child = c;
      c->prof_parent = ctx;
      c->check_sites = c->check_cells = 0;
    }
    jgm_context* c = sm.child;
    c->stream = ctx->stream;
    if (jgm_set_mesh(c, ctx->nx, ctx->ny, sm.nzv, ctx->xc.data(), ctx->yc.data(), zlo.data(),
                     ctx->xs.data(), ctx->ys.data(), zsz.data()) != JGM_OK ||
        jgm_set_sites(c, ctx->S, ctx->hsx.data(), ctx->hsy.data(), ctx->hsz.data()) != JGM_OK)
      return hipErrorInvalidValue;
    // the far field runs over the slabs only: the gap layers are empty in every model and
    // their gradient entries are not used (half the far-field pairs; the slabs all have one
    // shape on a mesh of uniform columns)
    c->far_layers = layer_of;
    c->near.valid = false;
    // sites with the site or its 1 m-lower point on a mesh line
    std::vector<int> line;
    for (size_t s = 0; s < ctx->S; ++s) {
      const double z = ctx->hsz[s];
      const bool zon = on_plane(Z, z) || on_plane(Z, z + 1.0);
      bool zslab = false;  // z on a slab plane Z_K - 1 as the slab mesh forms it
      for (size_t K = 0; K <= nz && !zslab; ++K) zslab = z == Z[K] - 1.0;
      const int on = (int)on_plane(ctx->xc, ctx->hsx[s]) + (int)on_plane(ctx->yc, ctx->hsy[s]) +
                     (int)(zon || zslab);
      if (on >= 2) line.push_back((int)s);
    }
    sm.nline = (int)line.size();
    hipError_t e;
    if ((e = upload_vec(sm.plane_of, plane_of, ctx->stream)) != hipSuccess ||
        (e = upload_vec(sm.layer_of, layer_of, ctx->stream)) != hipSuccess)
      return e;
    if (sm.nline > 0) {
      std::vector<double> xyz(3 * (size_t)sm.nline);
      for (int f = 0; f < sm.nline; ++f) {
        xyz[f] = ctx->hsx[line[f]];
        xyz[sm.nline + f] = ctx->hsy[line[f]];
        xyz[2 * (size_t)sm.nline + f] = ctx->hsz[line[f]];
      }
      if ((e = upload_vec(sm.line_idx, line, ctx->stream)) != hipSuccess ||
          (e = upload_vec(sm.line_xyz, xyz, ctx->stream)) != hipSuccess)
        return e;
    }
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    sm.ok = true;
  }
  // settings that may change without a new geometry
  jgm_context* c = sm.child;
  c->stream = ctx->stream;
  if (c->mixed_q0 != ctx->mixed_q0) {
    c->mixed_q0 = ctx->mixed_q0;
    c->near.valid = false;
  }
  c->mixed_q1 = ctx->mixed_q1;
  c->far_fwd_waves = ctx->far_fwd_waves;
  c->far_adj_waves = ctx->far_adj_waves;
  return hipSuccess;
}

}  // namespace

hipError_t launch_forward_maggrad_mixed(jgm_context* ctx, const double* model, double* out,
                                        double* raw_out) {
  // no cells: the FP64 path's zeros (nothing to expand)
  hipError_t e = ncells(ctx) == 0 ? hipErrorNotSupported : ensure_slab_mesh(ctx);
  if (e == hipErrorNotSupported)
    return launch_forward_fp64(ctx, JGM_KIND_MAGNETIC_GRADIENT, model, out, raw_out);
  if (e != hipSuccess) return e;
  SlabMesh& sm = ctx->slab;
  jgm_context* c = sm.child;
  const int S = (int)ctx->S;
  const long ncol = (long)ctx->nx * ctx->ny;
  const long Nv = ncol * (long)sm.nzv;
  if ((e = sm.rho.ensure(sizeof(double) * (size_t)std::max<long>(1, Nv))) != hipSuccess ||
      (e = sm.sums.ensure(sizeof(double) * 3 * (size_t)std::max(1, S))) != hipSuccess)
    return e;
  double bx, by, bz;
  field_unit(ctx, &bx, &by, &bz);
  {
    ProfScope ps(ctx, "maggrad_mixed_forward");
    hipLaunchKernelGGL(k_slab_weights, dim3(blocks_for(Nv)), dim3(kBlock), 0, ctx->stream, model,
                       sm.plane_of.as<const int>(), ncol, (int)ctx->nz, (int)sm.nzv,
                       sm.rho.as<double>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_zcol_sums_mixed(c, sm.rho.as<const double>(), sm.sums.as<double>())) !=
        hipSuccess)
      return e;
    hipLaunchKernelGGL(k_slab_finish, dim3(blocks_for(S)), dim3(kBlock), 0, ctx->stream,
                       sm.sums.as<const double>(), S, bx, by, bz, slab_scale(ctx), out, raw_out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (sm.nline > 0) {
    const double* xyz = sm.line_xyz.as<const double>();
    const int F = sm.nline;
    if ((e = launch_maggrad_fp64_sites(ctx, model, F, xyz, xyz + F, xyz + 2 * F,
                                       sm.line_idx.as<const int>(), out, raw_out)) != hipSuccess)
      return e;
  }
  return hipSuccess;
}

hipError_t launch_adjoint_maggrad_mixed(jgm_context* ctx, const double* rm, double* grad) {
  hipError_t e = ncells(ctx) == 0 ? hipErrorNotSupported : ensure_slab_mesh(ctx);
  if (e == hipErrorNotSupported)
    return launch_adjoint_fp64(ctx, JGM_KIND_MAGNETIC_GRADIENT, nullptr, rm, grad);
  if (e != hipSuccess) return e;
  SlabMesh& sm = ctx->slab;
  jgm_context* c = sm.child;
  const int S = (int)ctx->S;
  const long N = (long)ncells(ctx);
  const int L = (int)ctx->bgthick.size();
  const long ncol = (long)ctx->nx * ctx->ny;
  const long Nv = ncol * (long)sm.nzv;
  // OMPMagneticGradImp::CalcBackground returns zeros: no background gradient
  if (L > 0 && (e = launch_fill_zero(ctx, grad + N, (size_t)L)) != hipSuccess) return e;
  if (N == 0 || S == 0) return hipSuccess;
  if ((e = sm.rmt.ensure(sizeof(double) * 9 * (size_t)S)) != hipSuccess ||
      (e = sm.gv.ensure(sizeof(double) * (size_t)Nv)) != hipSuccess)
    return e;
  const double* r = rm;
  if (sm.nline > 0) {  // the line sites' misfits go through the FP64 form below
    if ((e = sm.rm0.ensure(sizeof(double) * (size_t)S)) != hipSuccess ||
        (e = hipMemcpyAsync(sm.rm0.ptr, rm, sizeof(double) * S, hipMemcpyDeviceToDevice,
                            ctx->stream)) != hipSuccess ||
        (e = launch_zero_site_rows(ctx, sm.rm0.as<double>(), sm.line_idx.as<const int>(), sm.nline,
                                   1)) != hipSuccess)
      return e;
    r = sm.rm0.as<const double>();
  }
  double bx, by, bz;
  field_unit(ctx, &bx, &by, &bz);
  {
    ProfScope ps(ctx, "maggrad_mixed_adjoint");
    hipLaunchKernelGGL(k_slab_tensor_misfit, dim3(blocks_for(9L * S)), dim3(kBlock), 0,
                       ctx->stream, r, S, bx, by, bz, slab_scale(ctx), sm.rmt.as<double>());
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_zcol_adjoint_mixed(c, sm.rmt.as<const double>(), sm.gv.as<double>())) !=
        hipSuccess)
      return e;
    hipLaunchKernelGGL(k_slab_gradient, dim3(blocks_for(N)), dim3(kBlock), 0, ctx->stream,
                       sm.gv.as<const double>(), sm.layer_of.as<const int>(), ncol, (int)ctx->nz,
                       (int)sm.nzv, grad);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (sm.nline > 0) {
    const double* xyz = sm.line_xyz.as<const double>();
    const int F = sm.nline;
    if ((e = launch_maggrad_fp64_sites_add(ctx, F, xyz, xyz + F, xyz + 2 * F,
                                           sm.line_idx.as<const int>(), rm, grad)) != hipSuccess)
      return e;
  }
  return hipSuccess;
}

bool maggrad_mixed_slabs(jgm_context* ctx) {
  return ensure_slab_mesh(ctx) == hipSuccess;
}

void release_slab_mesh(jgm_context* ctx) {
  if (ctx->slab.child) (void)jgm_destroy(ctx->slab.child);
  ctx->slab.child = nullptr;
  ctx->slab.gen = 0;
}

}  // namespace jgm
