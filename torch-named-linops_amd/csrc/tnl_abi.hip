// extern "C" entry points for interpolation / gridding and error plumbing.
#include <cstring>
#include <mutex>
#include <vector>

#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return TNL_OK;
  return fail(TNL_EHIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

// Validates the arguments shared by tnl_ungrid / tnl_grid and fills the
// launch geometry.  Mirrors the host-side checks of the reference:
// unknown kernel (kernels.py:97-105), norm not in {1, 2} (kernels.py:137),
// block width above 64 (grid.py:145-158).
static int make_geom(int ndim, const int64_t* grid_size, const float* width,
                     int kernel, int norm, int pad_mode, int is_complex,
                     float beta, InterpGeom& G) {
  if (ndim < 1 || ndim > 3) return fail(TNL_EINVAL, "ndim must be 1, 2 or 3");
  if (kernel != TNL_KERNEL_KAISER_BESSEL && kernel != TNL_KERNEL_SPLINE)
    return fail(TNL_EINVAL, "Unrecognized kernel");
  if (norm != 1 && norm != 2) return fail(TNL_EINVAL, "Unexpected norm");
  if (pad_mode != TNL_PAD_CIRCULAR && pad_mode != TNL_PAD_ZERO)
    return fail(TNL_EINVAL, "Unrecognized pad_mode");
  if (!grid_size || !width) return fail(TNL_EINVAL, "null size pointer");
  std::memset(&G, 0, sizeof(G));
  G.ndim = ndim;
  G.kernel = kernel;
  G.norm = norm;
  G.pad = pad_mode;
  G.beta = beta;
  for (int d = 0; d < 3; ++d) {
    G.n[d] = 1;
    G.half[d] = 0.5f;
    G.inv_half[d] = 2.0f;
    G.ntap[d] = 1;
  }
  for (int d = 0; d < ndim; ++d) {
    const int gd = 3 - ndim + d;
    if (grid_size[d] < 1) return fail(TNL_EINVAL, "grid_size must be >= 1");
    const float w = width[d];
    if (!(w > 0.0f) || !std::isfinite(w))
      return fail(TNL_EINVAL, "kernel width must be positive and finite");
    // reference: next_pow2(ceil(w + 1)) <= 64, last dim doubled if complex
    int64_t bw = 1;
    while (bw < (int64_t)std::ceil((double)w + 1.0)) bw *= 2;
    if (is_complex && d == ndim - 1) bw *= 2;
    if (bw > TNL_MAX_TAPS)
      return fail(TNL_EINVAL,
                  "Necessary block width exceeds maximum width 64 (kernel "
                  "width is doubled in last dim if input is complex)");
    G.n[gd] = grid_size[d];
    G.half[gd] = w * 0.5f;
    G.inv_half[gd] = 1.0f / (w * 0.5f);
    G.ntap[gd] = (int)std::ceil((double)w) + 1;
  }
  G.cstride[2] = 1;
  G.cstride[1] = G.n[2];
  G.cstride[0] = G.n[1] * G.n[2];
  G.ncells = G.n[0] * G.n[1] * G.n[2];
  return TNL_OK;
}

static int interp_entry(bool adjoint, const float* vals, const float* locs,
                        float* out, int64_t nbatch, int64_t npts, int ndim,
                        const int64_t* grid_size, const float* width,
                        int kernel, int norm, int pad_mode, int is_complex,
                        float beta, const tnl_plan_layout* layout,
                        const void* plan, void* scratch, int64_t scratch_bytes,
                        void* stream) {
  InterpGeom G;
  if (int rc = make_geom(ndim, grid_size, width, kernel, norm, pad_mode,
                         is_complex, beta, G))
    return rc;
  if (nbatch < 0 || npts < 0)
    return fail(TNL_EINVAL, "nbatch and npts must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const int64_t ncomp = is_complex ? 2 : 1;
  if (nbatch == 0) return TNL_OK;
  if (!out) return fail(TNL_EINVAL, "null output pointer");
  if (adjoint) {
    if (npts == 0) return fill_zero(out, nbatch * G.ncells * ncomp, s);
  } else if (npts == 0) {
    return TNL_OK;
  }
  if (!vals || !locs) return fail(TNL_EINVAL, "null input pointer");
  if ((plan == nullptr) != (layout == nullptr))
    return fail(TNL_EINVAL, "pass both the plan layout and its workspace, or neither");
  if (plan) {
    const tnl_plan_layout& L = *layout;
    bool same = L.ndim == ndim && L.npts == npts && L.pad_mode == pad_mode;
    for (int d = 0; same && d < 3; ++d)
      same = L.grid_size[d] == G.n[d] && L.ntaps[d] == G.ntap[d] &&
             L.width[d] == (d >= 3 - ndim ? width[d - (3 - ndim)] : 1.0f);
    if (!same)
      return fail(TNL_EINVAL,
                  "plan was built for another trajectory size, grid, width or pad_mode");
    if (binned_supported(G, L, adjoint)) {
      if (scratch && adjoint) {
        if (scratch_bytes < grid_scratch_bytes(L, nbatch, is_complex))
          return fail(TNL_EINVAL, "scratch smaller than tnl_grid_scratch_bytes");
      }
      if (scratch && !adjoint) {
        if (scratch_bytes < npts * nbatch * ncomp * (int64_t)sizeof(float))
          return fail(TNL_EINVAL, "scratch smaller than tnl_ungrid_scratch_bytes");
        if (reinterpret_cast<uintptr_t>(scratch) % 16 != 0)
          return fail(TNL_EINVAL, "scratch must be 16-byte aligned");
      }
      return interp_binned(adjoint, vals, locs, out, nbatch, G, is_complex, L, plan,
                           (float*)scratch, s);
    }
  }
  if (adjoint) {
    if (int rc = fill_zero(out, nbatch * G.ncells * ncomp, s)) return rc;
  }
  return interp_direct(adjoint, vals, locs, out, nbatch, npts, G, is_complex,
                       s);
}

}  // namespace tnl

using namespace tnl;

extern "C" const char* tnl_last_error(void) { return g_last_error.c_str(); }

extern "C" int tnl_abi_version(void) { return 1; }

extern "C" int tnl_ungrid(const float* vals, const float* locs, float* out,
                          int64_t nbatch, int64_t npts, int ndim,
                          const int64_t* grid_size, const float* width,
                          int kernel, int norm, int pad_mode, int is_complex,
                          float beta, const tnl_plan_layout* layout,
                          const void* plan, void* stream) {
  return interp_entry(false, vals, locs, out, nbatch, npts, ndim, grid_size,
                      width, kernel, norm, pad_mode, is_complex, beta, layout,
                      plan, nullptr, 0, stream);
}

extern "C" int tnl_grid(const float* vals, const float* locs, float* out,
                        int64_t nbatch, int64_t npts, int ndim,
                        const int64_t* grid_size, const float* width,
                        int kernel, int norm, int pad_mode, int is_complex,
                        float beta, const tnl_plan_layout* layout,
                        const void* plan, void* stream) {
  return interp_entry(true, vals, locs, out, nbatch, npts, ndim, grid_size,
                      width, kernel, norm, pad_mode, is_complex, beta, layout,
                      plan, nullptr, 0, stream);
}

extern "C" int tnl_ungrid_scratch_bytes(int64_t npts, int64_t nbatch, int is_complex,
                                        int64_t* bytes) {
  if (!bytes) return fail(TNL_EINVAL, "null pointer");
  if (npts < 0 || nbatch < 0) return fail(TNL_EINVAL, "npts and nbatch must be >= 0");
  *bytes = npts * nbatch * (is_complex ? 2 : 1) * (int64_t)sizeof(float);
  return TNL_OK;
}

extern "C" int tnl_ungrid_ws(const float* vals, const float* locs, float* out,
                             int64_t nbatch, int64_t npts, int ndim,
                             const int64_t* grid_size, const float* width,
                             int kernel, int norm, int pad_mode, int is_complex,
                             float beta, const tnl_plan_layout* layout,
                             const void* plan, void* scratch, int64_t scratch_bytes,
                             void* stream) {
  if (scratch_bytes < 0) return fail(TNL_EINVAL, "scratch_bytes must be >= 0");
  return interp_entry(false, vals, locs, out, nbatch, npts, ndim, grid_size,
                      width, kernel, norm, pad_mode, is_complex, beta, layout,
                      plan, scratch, scratch_bytes, stream);
}

extern "C" int tnl_grid_scratch_bytes(const tnl_plan_layout* layout, int64_t nbatch,
                                      int is_complex, int64_t* bytes) {
  if (!layout || !bytes) return fail(TNL_EINVAL, "null pointer");
  if (nbatch < 0) return fail(TNL_EINVAL, "nbatch must be >= 0");
  *bytes = grid_scratch_bytes(*layout, nbatch, is_complex);
  return TNL_OK;
}

extern "C" int tnl_grid_ws(const float* vals, const float* locs, float* out,
                           int64_t nbatch, int64_t npts, int ndim,
                           const int64_t* grid_size, const float* width,
                           int kernel, int norm, int pad_mode, int is_complex,
                           float beta, const tnl_plan_layout* layout,
                           const void* plan, void* scratch, int64_t scratch_bytes,
                           void* stream) {
  if (scratch_bytes < 0) return fail(TNL_EINVAL, "scratch_bytes must be >= 0");
  return interp_entry(true, vals, locs, out, nbatch, npts, ndim, grid_size,
                      width, kernel, norm, pad_mode, is_complex, beta, layout,
                      plan, scratch_bytes > 0 ? scratch : nullptr, scratch_bytes, stream);
}

// ---- float64 / complex128 interpolation (plan-free kernels, double arithmetic)
static int interp_f64_entry(bool adjoint, const double* vals, const double* locs, double* out,
                            int64_t nbatch, int64_t npts, int ndim, const int64_t* grid_size,
                            const float* width, int kernel, int norm, int pad_mode,
                            int is_complex, double beta, void* stream) {
  InterpGeom G;
  if (int rc = make_geom(ndim, grid_size, width, kernel, norm, pad_mode, is_complex,
                         (float)beta, G))
    return rc;
  if (nbatch < 0 || npts < 0) return fail(TNL_EINVAL, "nbatch and npts must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  const int64_t ncomp = is_complex ? 2 : 1;
  if (nbatch == 0) return TNL_OK;
  if (!out) return fail(TNL_EINVAL, "null output pointer");
  if (adjoint) {
    // zero bits are 0.0 in double too: fill two floats per double
    if (int rc = fill_zero(reinterpret_cast<float*>(out), 2 * nbatch * G.ncells * ncomp, s))
      return rc;
  }
  if (npts == 0) return TNL_OK;
  if (!vals || !locs) return fail(TNL_EINVAL, "null input pointer");
  return interp_direct_f64(adjoint, vals, locs, out, nbatch, npts, G, is_complex, beta, s);
}

extern "C" int tnl_ungrid_f64(const double* vals, const double* locs, double* out,
                              int64_t nbatch, int64_t npts, int ndim,
                              const int64_t* grid_size, const float* width, int kernel,
                              int norm, int pad_mode, int is_complex, double beta,
                              void* stream) {
  return interp_f64_entry(false, vals, locs, out, nbatch, npts, ndim, grid_size, width,
                          kernel, norm, pad_mode, is_complex, beta, stream);
}

extern "C" int tnl_grid_f64(const double* vals, const double* locs, double* out,
                            int64_t nbatch, int64_t npts, int ndim, const int64_t* grid_size,
                            const float* width, int kernel, int norm, int pad_mode,
                            int is_complex, double beta, void* stream) {
  return interp_f64_entry(true, vals, locs, out, nbatch, npts, ndim, grid_size, width,
                          kernel, norm, pad_mode, is_complex, beta, stream);
}

// ---- allocation-free gridding: the point records in a static device array
// Each GPU holds its own copy of `scratch` at its own address; host code
// never passes the array's name to a kernel (that host-side name is not a
// device address) but the per-device pointer hipGetSymbolAddress returns,
// taken once per device with that device current.
static __device__ __align__(256) unsigned char scratch[TNL_GRID_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static int static_scratch(void** ptr) {
  int dev = 0;
  if (int rc = check_hip(hipGetDevice(&dev), "hipGetDevice")) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail(TNL_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (int rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(scratch)"))
      return rc;
    if (!p) return fail(TNL_EHIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return TNL_OK;
}

extern "C" int tnl_grid_static(const float* vals, const float* locs, float* out,
                               int64_t nbatch, int64_t npts, int ndim,
                               const int64_t* grid_size, const float* width,
                               int kernel, int norm, int pad_mode, int is_complex,
                               float beta, const tnl_plan_layout* layout,
                               const void* plan, void* stream) {
  if (nbatch < 0) return fail(TNL_EINVAL, "nbatch must be >= 0");
  const int64_t need = layout ? grid_scratch_bytes(*layout, nbatch, is_complex) : 0;
  if (need > TNL_GRID_STATIC_SCRATCH_BYTES)
    return fail(TNL_EINVAL, "tnl_grid_static: this plan needs " + std::to_string(need) +
                                " bytes of scratch, more than the static " +
                                std::to_string((long long)TNL_GRID_STATIC_SCRATCH_BYTES) +
                                " bytes; use tnl_grid_ws with a caller workspace");
  void* p = nullptr;
  if (need > 0 && plan && nbatch > 0 && npts > 0) {
    if (int rc = static_scratch(&p)) return rc;
  }
  return interp_entry(true, vals, locs, out, nbatch, npts, ndim, grid_size, width, kernel,
                      norm, pad_mode, is_complex, beta, layout, plan, p, p ? need : 0,
                      stream);
}

extern "C" int tnl_interp_grad_locs(const float* grid_vals, const float* pt_vals,
                                    const float* locs, float* grad_locs,
                                    int64_t nbatch, int64_t npts, int ndim,
                                    const int64_t* grid_size, const float* width,
                                    int kernel, int norm, int pad_mode,
                                    int is_complex, float beta, void* stream) {
  InterpGeom G;
  if (int rc = make_geom(ndim, grid_size, width, kernel, norm, pad_mode,
                         is_complex, beta, G))
    return rc;
  if (nbatch < 0 || npts < 0)
    return fail(TNL_EINVAL, "nbatch and npts must be >= 0");
  if (npts == 0) return TNL_OK;
  if (!grad_locs || !locs) return fail(TNL_EINVAL, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (nbatch == 0) return fill_zero(grad_locs, npts * ndim, s);
  if (!grid_vals || !pt_vals) return fail(TNL_EINVAL, "null input pointer");
  return interp_grad_locs(grid_vals, pt_vals, locs, grad_locs, nbatch, npts, G,
                          is_complex, s);
}

// ------------------------------------------------------------- batches
static int batch_geoms(const tnl_batch_problem* probs, int nprob, int ndim,
                       int kernel, int norm, int pad_mode, int is_complex,
                       InterpGeom* G) {
  for (int i = 0; i < nprob; ++i) {
    const tnl_batch_problem& p = probs[i];
    float width[3] = {p.width[0], p.width[1], p.width[2]};
    if (int rc = make_geom(ndim, p.grid_size, width, kernel, norm, pad_mode,
                           is_complex, p.beta, G[i]))
      return rc;
    if (p.nbatch < 1 || p.npts < 0 || p.grid_offset < 0 || p.pts_offset < 0)
      return fail(TNL_EINVAL, "batch problem: nbatch >= 1, npts and offsets >= 0");
    if (!p.layout || !p.plan) return fail(TNL_EINVAL, "batch problem without a plan");
    const tnl_plan_layout& L = *p.layout;
    bool same = L.ndim == ndim && L.npts == p.npts && L.pad_mode == pad_mode;
    for (int d = 0; same && d < 3; ++d)
      same = L.grid_size[d] == G[i].n[d] && L.ntaps[d] == G[i].ntap[d] &&
             L.width[d] == (d >= 3 - ndim ? width[d - (3 - ndim)] : 1.0f);
    if (!same)
      return fail(TNL_EINVAL, "batch problem: plan built for another size, grid, width or pad_mode");
  }
  return TNL_OK;
}

extern "C" int tnl_batch_layout_query(const tnl_batch_problem* probs, int nprob,
                                      int ndim, int kernel, int norm,
                                      int pad_mode, int is_complex,
                                      tnl_batch_layout* layout) {
  if (!probs || !layout || nprob < 1)
    return fail(TNL_EINVAL, "batch needs at least one problem");
  std::vector<InterpGeom> G((size_t)nprob);
  if (int rc = batch_geoms(probs, nprob, ndim, kernel, norm, pad_mode, is_complex, G.data()))
    return rc;
  std::memset(layout, 0, sizeof(*layout));
  layout->nprob = nprob;
  layout->ndim = ndim;
  layout->kernel = kernel;
  layout->norm = norm;
  layout->pad_mode = pad_mode;
  layout->is_complex = is_complex ? 1 : 0;
  return batch_prepare(G.data(), probs, nprob, is_complex, layout);
}

extern "C" int tnl_batch_build(const tnl_batch_problem* probs,
                               const tnl_batch_layout* layout, void* ws,
                               void* stream) {
  if (!probs || !layout || !ws) return fail(TNL_EINVAL, "null pointer");
  std::vector<InterpGeom> G((size_t)layout->nprob);
  if (int rc = batch_geoms(probs, layout->nprob, layout->ndim, layout->kernel,
                           layout->norm, layout->pad_mode, layout->is_complex, G.data()))
    return rc;
  return batch_build(G.data(), probs, *layout, ws, (hipStream_t)stream);
}

extern "C" int tnl_ungrid_batched(const float* grid_vals, float* pts_out,
                                  const tnl_batch_layout* layout,
                                  const void* ws, void* stream) {
  if (!grid_vals || !pts_out || !layout || !ws) return fail(TNL_EINVAL, "null pointer");
  return batch_run(false, *layout, ws, grid_vals, pts_out, (hipStream_t)stream);
}

extern "C" int tnl_grid_batched(const float* pts_vals, float* grid_out,
                                const tnl_batch_layout* layout, const void* ws,
                                void* stream) {
  if (!pts_vals || !grid_out || !layout || !ws) return fail(TNL_EINVAL, "null pointer");
  return batch_run(true, *layout, ws, pts_vals, grid_out, (hipStream_t)stream);
}
