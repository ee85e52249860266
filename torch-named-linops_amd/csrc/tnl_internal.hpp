// Host-side declarations shared by the libtnl translation units.
#pragma once

#include "tnl_common.hpp"

namespace tnl {

int fill_zero(float* p, int64_t n, hipStream_t s);
int interp_direct_f64(bool adjoint, const double* vals, const double* locs, double* out,
                      int64_t nbatch, int64_t npts, const InterpGeom& G, int is_complex,
                      double beta, hipStream_t s);

int interp_direct(bool adjoint, const float* vals, const float* locs,
                  float* out, int64_t nbatch, int64_t npts,
                  const InterpGeom& G, int is_complex, hipStream_t s);

// Binned (plan-driven) path; returns TNL_EINVAL with no launch when the
// problem is outside what the binned kernels cover (caller then falls back).
bool binned_supported(const InterpGeom& G, const tnl_plan_layout& L,
                      bool adjoint);
// scratch: NULL, or grid_scratch_bytes(L, nbatch, is_complex) bytes for the
// point-major records of matrix-core gridding (ignored elsewhere).
int interp_binned(bool adjoint, const float* vals, const float* locs, float* out,
                  int64_t nbatch, const InterpGeom& G, int is_complex,
                  const tnl_plan_layout& L, const void* plan, float* scratch,
                  hipStream_t s);
int64_t grid_scratch_bytes(const tnl_plan_layout& L, int64_t nbatch, int is_complex);

// One-wave matrix-core gridding (tnl_gridwave.hip): plans with touch pairs.
bool grid_wave_ok(const tnl_plan_layout& L);
int64_t grid_wave_wrec_bytes(const tnl_plan_layout& L);
// rec: the point-major channel records (filled by the caller), wrec:
// grid_wave_wrec_bytes of scratch for the window records (filled here).
int grid_wave_launch(const float* vals, const float* locs, float* out, int64_t nbatch,
                     const InterpGeom& G, const tnl_plan_layout& L, const char* ws,
                     const float* rec, float* wrec, int is_complex, hipStream_t s);

int interp_grad_locs(const float* grid_vals, const float* pt_vals,
                     const float* locs, float* grad, int64_t nbatch,
                     int64_t npts, const InterpGeom& G, int is_complex,
                     hipStream_t s);

int batch_prepare(const InterpGeom* G, const tnl_batch_problem* P, int nprob,
                  int is_complex, tnl_batch_layout* L);
int batch_build(const InterpGeom* G, const tnl_batch_problem* P,
                const tnl_batch_layout& L, void* ws, hipStream_t s);
int batch_run(bool adjoint, const tnl_batch_layout& L, const void* ws,
              const float* in, float* out, hipStream_t s);

int seg_layout(const int64_t* seg_len, int nseg, int64_t block_floats,
               tnl_seg_layout* L);
int seg_build(const int64_t* seg_len, const tnl_seg_layout& L, void* ws,
              hipStream_t s);
int seg_dot(const float* a, const float* b, double* out, const tnl_seg_layout& L,
            void* ws, hipStream_t s);
int seg_cg_update(float* x, float* r, const float* p, const float* Ap,
                  const double* rs, const double* pAp, double* rs_new,
                  const tnl_seg_layout& L, void* ws, hipStream_t s);
int seg_cg_direction(float* p, const float* r, double* rs, const double* rs_new,
                     const tnl_seg_layout& L, void* ws, hipStream_t s);

int conv3d_fwd(const void* in, const void* w, const float* bias, void* out_b,
               float* out_f, int64_t nbatch, int Z, int Y, int X, int cin_p,
               int cout, int cout_p, int relu, hipStream_t s);
int64_t conv3d_wgrad_ws_bytes(int64_t nbatch, int Z, int Y, int X);
int conv3d_wgrad(const void* in, const void* dout, float* dw, float* ws,
                 int64_t ws_bytes, int64_t nbatch, int Z, int Y, int X, int cin_p,
                 int cout, int cout_p, hipStream_t s);

int plan_layout(int64_t npts, int ndim, const int64_t* grid_size,
                const float* width, int pad_mode, const tnl_tuning* tune,
                tnl_plan_layout* L);

}  // namespace tnl
