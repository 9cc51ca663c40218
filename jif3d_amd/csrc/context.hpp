// Internal state behind the C-ABI handle (include/jif3d_gravmag.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "jif3d_gravmag.h"

namespace jgm {

// Grow-only device allocation owned by a context; freed by release() or on
// destruction, never copied.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (ptr) {
      hipError_t e = hipFree(ptr);
      if (e != hipSuccess) return e;
      ptr = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&ptr, need < 256 ? 256 : need);
    if (e == hipSuccess) bytes = need < 256 ? 256 : need;
    return e;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Near-field description for the mixed-precision kernels: per site an index box
// of cells [lo, hi] (inclusive) evaluated with the FP64 closed form; every cell
// outside it is at least q0 of its own layer's largest half-diagonals away from the site.
struct NearBoxes {
  DeviceBuffer box;       // int32 [S][6]: i0, i1, j0, j1, k0, k1 (empty: i0 > i1)
  DeviceBuffer col_ptr;   // int32 [ncols+1]: CSR of sites whose box covers column (i, j)
  DeviceBuffer col_sites; // int32 site indices, ascending within a column
  DeviceBuffer col_ids;   // int32 column index i*ny + j of each listed column
  DeviceBuffer col_kmax;  // int32 largest k1 over the column's sites
  DeviceBuffer corner_off; // int64 per-site offset of its box corners (adjoint near pass)
  DeviceBuffer cell_off;   // int64 per-site offset of its box cells (adjoint near pass)
  size_t ncols = 0;
  size_t total_near_cells = 0;    // sum over sites of box cells
  size_t total_near_corners = 0;  // sum over sites of box corners
  double radius = 0.0;          // q0 * largest cell half-diagonal of the mesh (m)
  bool valid = false;
  // far-field geometry: cell centres (FP64)
  DeviceBuffer xcen, ycen, zcen;
  // FP32 offset of each cell centre from the centre of the first cell of its 4-cell
  // tile run along the axis (exact for uniform meshes)
  DeviceBuffer xoff, yoff, zoff;
  // squared largest half-size of each 4-run of cells per axis (z also per 16-run): the
  // tiles' half-diagonals that scale the second radius
  DeviceBuffer hx2, hy2, hz2, hz2_16;
  int mesh_class = 2;  // 0: cubes of one size, 1: one cell shape, 2: general
  // the layers the far field runs over (jgm_context::far_layers; all when that is empty):
  // nzf of them, kmap (int32 [nzf]) their mesh layers in increasing order, listed (int32
  // [nz]) 1 for a listed layer, far_dz the cell height of the first of them.  The z arrays
  // above (zcen, zoff, hz2, hz2_16) run over the listed layers.
  DeviceBuffer kmap, listed;
  int nzf = 0;
  bool far_subset = false;
  double far_dz = 0.0;
  DeviceBuffer adj_table;  // FP32 per-cell table with unit density (G V, B...)
  bool adj_table_valid = false;
  DeviceBuffer site_order;  // int32 [S]: sites in Morton order of (x, y), for compact waves
  // cost-balanced partition of the far adjoint's work (geometry, q1 and the wave target
  // only): per workgroup of four adjoint tiles, adj_chunks + 1 site boundaries (int32)
  DeviceBuffer adj_ranges;
  int adj_chunks = 0;
  double plan_q1 = -1.0;
  bool plan_valid = false;
  // sites on mesh lines (two coordinates equal to corner coordinates of the mesh): the
  // mixed path evaluates them with the FP64 closed form; nline of them, their survey
  // indices (int32) and coordinates (FP64 x[nline], y[nline], z[nline])
  int nline = 0;
  DeviceBuffer line_idx, line_xyz;
  std::vector<int> line_sites;  // host copy of the indices
};

// Mixed-precision vertical magnetic gradient: the slab mesh (1 m slabs just above every
// mesh plane, zero-weight gaps between them) held by a child context, made per geometry
// (gen); plane_of[kv]: the mesh plane of slab layer kv (-1 for a gap), layer_of[K]: the
// layer of plane K's slab; nline sites evaluated with the FP64 two-point form.
struct SlabMesh {
  jgm_context* child = nullptr;
  unsigned long gen = 0;
  bool ok = false;
  size_t nzv = 0;
  int nline = 0;
  DeviceBuffer plane_of, layer_of, line_idx, line_xyz;
  DeviceBuffer rho, sums, rmt, gv, rm0;
};

// Optional per-kernel timing: HIP events recorded on the context's stream around
// each launch, harvested lazily.
struct Profiler {
  struct Entry {
    std::string name;
    double total_ms = 0.0;
    size_t launches = 0;
  };
  struct Pending {
    size_t entry;
    hipEvent_t start, stop;
  };
  bool enabled = false;
  std::vector<Entry> entries;
  std::vector<Pending> pending;
};

}  // namespace jgm

struct jgm_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int cu_count = 256;

  // mesh (host copies)
  size_t nx = 0, ny = 0, nz = 0;
  std::vector<double> xc, yc, zc;  // corner coordinates, n+1 each
  std::vector<double> xs, ys, zs;  // cell sizes, n each
  double xwidth = 0, ywidth = 0, zwidth = 0;
  bool mesh_set = false;
  // device mesh: corner coordinates used by the corner-first kernels (n+1 each),
  // and the per-cell origin/size arrays used by the per-cell kernels.
  jgm::DeviceBuffer d_Xc, d_Yc, d_Zc;
  jgm::DeviceBuffer d_xlo, d_ylo, d_zlo, d_xsz, d_ysz, d_zsz;

  // background layers
  std::vector<double> bgthick;

  // sites
  size_t S = 0;
  std::vector<double> hsx, hsy, hsz;
  jgm::DeviceBuffer d_sx, d_sy, d_sz;
  bool sites_set = false;

  // magnetic inducing field (radians, radians, nT)
  double inclination = 0.0, declination = 0.0, fieldstrength = 0.0;

  // data transform
  int transform = JGM_TF_NONE;
  double tf_inc = 0.0, tf_dec = 0.0;

  // mixed precision near-field threshold in cell half-diagonals, and the radius (also
  // in half-diagonals) beyond which the far field keeps only monopole (and quadrupole)
  // terms; 0 keeps the hexadecapole everywhere
  double mixed_q0 = 8.0;
  double mixed_q1 = 28.0;
  // waves per far-field launch the work is cut for: forward (equal chunks of tiles) and
  // adjoint (planned chunks of sites); environment variables JIF3D_AMD_FAR_FORWARD_WAVES and
  // JIF3D_AMD_FAR_ADJOINT_WAVES at creation
  long far_fwd_waves = 65536;
  long far_adj_waves = 49152;
  // mesh layers the mixed far field runs over, increasing (empty: every layer).  Every other
  // layer must carry zero weight in every model: the far field skips it, and the mixed
  // adjoint returns zero for its cells.  The vertical gradient's slab mesh lists its slabs
  // (its gap layers are empty).  Changing it needs the near boxes made again.
  std::vector<int> far_layers;

  // a-posteriori check of JGM_MIXED: FP64 at check_sites sample sites after a forward and
  // at check_cells sample cells after an adjoint; a sampled error above check_tol (relative
  // to the survey maximum of each component, or to the largest gradient entry) recomputes
  // the call in FP64.  Sample indices are uploaded per geometry (check_gen).
  size_t check_sites = 4, check_cells = 16;
  double check_tol = 5e-7;
  size_t check_F = 0, check_FC = 0;
  unsigned long check_gen = 0;
  double check_last_fwd = 0.0, check_last_adj = 0.0;
  size_t check_fallbacks = 0;
  jgm::DeviceBuffer check_sidx, check_cidx, check_w, check_part;

  // incremented by every setter: caches record the generation they were made for
  unsigned long config_gen = 1;
  // incremented by the mesh and site setters only
  unsigned long geom_gen = 1;
  // the vertical gradient's slab mesh (mixed precision)
  jgm::SlabMesh slab;
  // a child context records its kernel timings into this one's profiler
  jgm_context* prof_parent = nullptr;
  // raw (untransformed) data of the last jgm_forward made with a data transform, and a
  // copy of its model: an adjoint at the same model takes the transform's Jacobian from
  // it instead of a second forward
  struct RawCache {
    bool valid = false;
    int kind = 0, prec = 0;
    unsigned long gen = 0;
    size_t nmodel = 0;
  } raw_cache;
  jgm::DeviceBuffer raw_cache_model, raw_cache_raw, raw_cache_flag;

  // caches
  jgm::DeviceBuffer bgT;  // [S][L][7] background box terms, valid when bg_valid
  bool bg_valid = false;
  jgm::NearBoxes near;

  // scratch
  jgm::DeviceBuffer weights, partial, raw, coef, hpart, hsum, model_stage, out_stage, misfit_stage,
      grad_stage, farpart, nearpart, coef32, coef_mono, near_h, near_v, vsites, partsum,
      chunk_scratch, mag_part;
  // FP64 evaluation of the sites the mixed path hands back (corner weights, partials,
  // sums, gathered misfits, coefficients, corner sums, a gradient, the residuals of the
  // other sites)
  jgm::DeviceBuffer rt_w, rt_part, rt_sum, rt_rm, rt_coef, rt_h, rt_tmp, rt_rm0;

  jgm::Profiler prof;

  // stored sensitivity matrix: FP32, column-major with leading dimension storedJ_ld
  // (rows padded to a multiple of 64) and storedJ_colpad columns (cells padded to a
  // multiple of 256); the padding holds zeros
  jgm::DeviceBuffer storedJ, storedJ_vec;
  int storedJ_kind = 0;
  bool storedJ_fp64 = false;  // J held in FP64 (jgm_storedj_build_fp64) instead of FP32
  size_t storedJ_rows = 0, storedJ_cols = 0, storedJ_ld = 0, storedJ_colpad = 0;
  // host-resident stored J: the whole J in page-locked host memory, streamed in column
  // blocks of storedJ_block cells through two HBM staging buffers (storedJ then holds
  // the staging buffers); null when J is in HBM
  float* storedJ_host = nullptr;
  size_t storedJ_block = 0;
  // file-backed stored J: the same layout in a file (storedJ_fd open, removed on release),
  // read block by block into two page-locked buffers and streamed like the host-resident J
  int storedJ_fd = -1;
  std::string storedJ_file;
  float* storedJ_pin[2] = {nullptr, nullptr};
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_used[2] = {nullptr, nullptr};
  hipEvent_t ev_start = nullptr;
  jgm::DeviceBuffer acc64;  // FP64 totals of J M over the streamed blocks
  // BF16 copy of an HBM-resident J for the BF16 J M, 8 cells packed per row; made from
  // the J of generation storedJ_bf16_gen (storedJ_gen counts builds).  A failed
  // allocation of the copy is remembered for the generation it failed on, so it is not
  // tried again until J is rebuilt; bf16_copy_enabled = false (environment variable
  // JIF3D_AMD_BF16_COPY=0 when the context is created) never makes the copy.
  jgm::DeviceBuffer storedJ_bf16;
  size_t storedJ_gen = 1, storedJ_bf16_gen = 0, storedJ_bf16_failed_gen = 0;
  bool bf16_copy_enabled = true;
  // BF16 copy of J in J's own column-major layout (a cell's rows contiguous) for the BF16
  // J^T R, made and dropped like the copy above (same switch)
  jgm::DeviceBuffer storedJ_bf16c, storedJ_rhs16;  // the copy; its BF16 right-hand sides
  size_t storedJ_bf16c_gen = 0, storedJ_bf16c_failed_gen = 0;
  // the low BF16 parts of the right-hand sides of the BF16-refined J M (storedJ_rhs16 the
  // high parts), made per call
  jgm::DeviceBuffer storedJ_rhs16lo;
};

namespace jgm {

// Launch helpers implemented in the kernel translation units.  All return a HIP
// error code; arguments have been validated by the API layer.
hipError_t launch_background_table(jgm_context* ctx);
hipError_t launch_forward_fp64(jgm_context* ctx, int kind, const double* model, double* out,
                               double* raw_out);
hipError_t launch_adjoint_fp64(jgm_context* ctx, int kind, const double* model,
                               const double* raw_misfit, double* grad);
// batches: nmodels models / residual sets, packed with the given leading dimensions
hipError_t launch_forward_fp64_batch(jgm_context* ctx, int kind, size_t nmodels,
                                     const double* models, size_t model_ld, double* out,
                                     size_t out_ld, double* raw_out, size_t raw_ld);
hipError_t launch_adjoint_fp64_batch(jgm_context* ctx, int kind, size_t nsets,
                                     const double* raw_misfits, size_t misfit_ld, double* grads,
                                     size_t grad_ld);
hipError_t launch_misfit_to_raw(jgm_context* ctx, int kind, const double* misfit,
                                const double* raw, double* raw_misfit);
hipError_t launch_sensitivity_rows(jgm_context* ctx, int kind, size_t site0, size_t nsites,
                                   double* rows);
hipError_t launch_forward_mixed(jgm_context* ctx, int kind, const double* model, double* out,
                                double* raw_out);
hipError_t launch_forward_mixed_batch(jgm_context* ctx, int kind, size_t nmodels,
                                      const double* models, size_t model_ld, double* out,
                                      size_t out_ld, double* raw_out, size_t raw_ld);
hipError_t launch_adjoint_mixed(jgm_context* ctx, int kind, const double* model,
                                const double* raw_misfit, double* grad);
// nmodels gradients, in groups of 4, 2 and 1 that share the far-field kernel terms
hipError_t launch_adjoint_mixed_batch(jgm_context* ctx, int kind, size_t nmodels,
                                      const double* raw_misfits, size_t rm_ld, double* grads,
                                      size_t grad_ld);
hipError_t build_near_boxes(jgm_context* ctx);
// allocation-free FP64 forward (fp64_path.hip): the device scratch array's address is
// looked up once per device (called by jgm_create with the device current);
// noalloc_forward_need gives the bytes of scratch one call needs with a single chunk
// of corner columns; the launch returns hipErrorOutOfMemory when even that does not fit
// FP64 closed form for F of the sites (coordinates fx, fy, fz, survey indices idx):
// the forward writes their data into out / raw_out at their own indices; the adjoint adds
// their cell gradients (rm: raw misfits of the whole survey) into grad
hipError_t launch_forward_fp64_sites(jgm_context* ctx, int kind, const double* model, int F,
                                     const double* fx, const double* fy, const double* fz,
                                     const int* idx, double* out, double* raw_out);
hipError_t launch_adjoint_fp64_sites_add(jgm_context* ctx, int kind, int F, const double* fx,
                                         const double* fy, const double* fz, const int* idx,
                                         const double* rm, double* grad);
hipError_t launch_zero_site_rows(jgm_context* ctx, double* a, const int* idx, int F, int k);
// the vertical-gradient kind's FP64 two-point form at F listed sites (coordinates fx, fy,
// fz; survey indices idx): rows idx of out / raw_out, or gradients added into grad
hipError_t launch_maggrad_fp64_sites(jgm_context* ctx, const double* model, int F,
                                     const double* fx, const double* fy, const double* fz,
                                     const int* idx, double* out, double* raw_out);
hipError_t launch_maggrad_fp64_sites_add(jgm_context* ctx, int F, const double* fx,
                                         const double* fy, const double* fz, const int* idx,
                                         const double* rm, double* grad);
// mixed sums of the tensor's z column (S x 3: zz, xz, yz, carrying G and the signs) of one
// density model, and their mixed gradient for a raw tensor misfit (9 per site, only its
// zx, zy, zz entries reaching them); meshes of cubes excepted (hipErrorNotSupported)
hipError_t launch_zcol_sums_mixed(jgm_context* ctx, const double* rho, double* sums);
hipError_t launch_zcol_adjoint_mixed(jgm_context* ctx, const double* rmt, double* grad);
// the vertical-gradient kind in mixed precision through its slab mesh (maggrad_mixed.hip);
// FP64 when a cell is thinner than 1 m (maggrad_mixed_slabs false)
hipError_t launch_forward_maggrad_mixed(jgm_context* ctx, const double* model, double* out,
                                        double* raw_out);
hipError_t launch_adjoint_maggrad_mixed(jgm_context* ctx, const double* rm, double* grad);
bool maggrad_mixed_slabs(jgm_context* ctx);
void release_slab_mesh(jgm_context* ctx);
// a-posteriori check of a mixed forward (raw data of every site in mixed_raw) or adjoint
// (its gradient): the largest sampled error, relative as stated, in *err_host (waits)
bool mixed_check_applies(const jgm_context* ctx, int kind);
hipError_t launch_mixed_check_forward(jgm_context* ctx, int kind, const double* model,
                                      const double* mixed_raw, double* err_host);
hipError_t launch_mixed_check_adjoint(jgm_context* ctx, int kind, const double* raw_misfit,
                                      const double* grad, double* err_host);
void noalloc_scratch_init(int device);
double* noalloc_scratch(int device, hipError_t* err);
size_t noalloc_forward_need(const jgm_context* ctx, int kind);
hipError_t launch_forward_fp64_noalloc(jgm_context* ctx, int kind, const double* model,
                                       double* out);
hipError_t launch_fill_zero(jgm_context* ctx, double* p, size_t n);
// out[i] = sum over c < nchunks of part[c * n + i] (fixed order, two passes for many chunks)
hipError_t launch_reduce_chunks(jgm_context* ctx, const double* part, int nchunks, long n,
                                double* out);
// flag[0] = 0 if a[0..n) and b[0..n) are bitwise equal, 1 otherwise
hipError_t launch_models_differ(jgm_context* ctx, const double* a, const double* b, size_t n,
                                int* flag);
hipError_t launch_background_gradient(jgm_context* ctx, int kind, const double* raw_misfit,
                                      double* grad_bg);
hipError_t launch_storedj_build(jgm_context* ctx, int kind, int precision);
hipError_t launch_storedj_build_resident(jgm_context* ctx, int kind, int precision,
                                         size_t block_cells);
void release_host_storedj(jgm_context* ctx);
hipError_t launch_storedj_build_file(jgm_context* ctx, int kind, size_t block_cells,
                                     const char* dir);
hipError_t launch_storedj_apply(jgm_context* ctx, int gemm_precision, int transpose,
                                const float* in, size_t nrhs, float* out);
hipError_t launch_storedj_build_f64(jgm_context* ctx, int kind);
hipError_t launch_storedj_apply_f64(jgm_context* ctx, int transpose, const double* in,
                                    size_t nrhs, double* out);
hipError_t launch_storedj_forward(jgm_context* ctx, int gemm_precision, const double* model,
                                  double* out, double* raw_out);
hipError_t launch_storedj_gradient(jgm_context* ctx, int gemm_precision,
                                   const double* raw_misfit, double* grad);
hipError_t launch_storedj_forward_batch(jgm_context* ctx, int gemm_precision, size_t nmodels,
                                        const double* models, size_t model_ld, double* out,
                                        size_t out_ld, double* raw_out, size_t raw_ld);
hipError_t launch_storedj_gradient_batch(jgm_context* ctx, int gemm_precision, size_t nsets,
                                         const double* raw_misfits, size_t misfit_ld,
                                         double* grads, size_t grad_ld);

__host__ __device__ inline size_t raw_per_site(int kind) {
  switch (kind) {
    case JGM_KIND_SCALAR: return 1;
    case JGM_KIND_TENSOR: return 9;
    case JGM_KIND_MAGNETIC: return 3;
    case JGM_KIND_SCALAR_TENSOR: return 10;
    case JGM_KIND_MAGNETIC_GRADIENT: return 1;
    case JGM_KIND_MAGNETIZATION: return 3;
    default: return 0;
  }
}

// Kernel timing hooks (profiler.cpp).  prof_begin returns a token for prof_end,
// or -1 when profiling is off.
long prof_begin(jgm_context* ctx, const char* name);
void prof_end(jgm_context* ctx, long token);
hipError_t prof_harvest(jgm_context* ctx);
void prof_clear(jgm_context* ctx);

struct ProfScope {
  jgm_context* ctx;
  long token;
  ProfScope(jgm_context* c, const char* name) : ctx(c), token(prof_begin(c, name)) {}
  ~ProfScope() { prof_end(ctx, token); }
};

inline size_t ncells(const jgm_context* c) { return c->nx * c->ny * c->nz; }
inline size_t ncorners(const jgm_context* c) { return (c->nx + 1) * (c->ny + 1) * (c->nz + 1); }
inline size_t nparams(const jgm_context* c) { return ncells(c) + c->bgthick.size(); }
// model vector length of a kind: three values per cell and no background layers for
// the magnetization kind (ThreeDMagnetizationModel), cells then layers otherwise
inline size_t nparams(const jgm_context* c, int kind) {
  return kind == JGM_KIND_MAGNETIZATION ? 3 * ncells(c) : nparams(c);
}

}  // namespace jgm
