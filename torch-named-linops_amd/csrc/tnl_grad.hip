// Gradient of interpolation / gridding with respect to the trajectory.
//
// With W(g, t) the separable (norm 1) or radial (norm 2) window weight of
// grid point g for a sample at t, both
//   ungrid:  y_b[p]  = sum_g x_b[g] W(g, t_p)
//   grid:    X_b[g] += v_b[p] W(g, t_p)
// have trajectory gradients of the form
//   dL/dt_{p,d} = sum_b sum_g Re(conj(h_b[p]) x_b[g]) dW(g, t_p)/dt_{p,d}
// (ungrid: x = grid input, h = dL/dy; grid: x = dL/dX, h = v), so one
// kernel serves both.  One thread per sample gathers the grid values of its
// footprint; the window derivatives are analytic (kb_window_deriv).
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

constexpr int GRAD_ZT = 16;  // tap capacity per axis

template <bool CPLX>
struct GVal;
template <>
struct GVal<false> {
  using T = float;
  __device__ static T load(const float* p, int64_t i) { return p[i]; }
  // Re(conj(h) x)
  __device__ static float rdot(T h, T x) { return h * x; }
};
template <>
struct GVal<true> {
  using T = float2;
  __device__ static T load(const float* p, int64_t i) {
    return reinterpret_cast<const float2*>(p)[i];
  }
  __device__ static float rdot(T h, T x) { return fmaf(h.x, x.x, h.y * x.y); }
};

template <int NDIM, bool CPLX>
__global__ __launch_bounds__(256) void grad_locs_kernel(
    const float* __restrict__ gvals, const float* __restrict__ pvals,
    const float* __restrict__ locs, float* __restrict__ grad, int64_t nbatch,
    int64_t npts, InterpGeom G) {
  using V = GVal<CPLX>;
  constexpr int NB = 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npts;
       p += stride) {
    float t[3] = {0.f, 0.f, 0.f};
    TapRange r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      r[d].g0 = 0;
      r[d].count = 1;
      if (d >= 3 - NDIM) {
        t[d] = locs[p * NDIM + d - (3 - NDIM)];
        r[d] = tap_range(t[d], G.half[d], G.ntap[d], G.n[d], G.pad);
      }
    }
    // Per tap of one axis: u = |g - t| / (w/2), its derivative
    // du/dt = sign(t - g) / (w/2), the window and the window derivative
    // along t.  The z axis is cached; x and y are recomputed in the loops.
    float zu[GRAD_ZT], zs[GRAD_ZT], zw[GRAD_ZT], zd[GRAD_ZT];
    int64_t zo[GRAD_ZT];
#pragma unroll
    for (int k = 0; k < GRAD_ZT; ++k) {
      zu[k] = zs[k] = zw[k] = zd[k] = 0.f;
      zo[k] = 0;
      if (k < r[2].count) {
        const int64_t g = r[2].g0 + k;
        const float diff = t[2] - (float)g;
        zu[k] = fminf(fabsf(diff) * G.inv_half[2], 1.0f);
        zs[k] = (diff > 0.f ? 1.f : (diff < 0.f ? -1.f : 0.f)) * G.inv_half[2];
        zw[k] = window(G.kernel, zu[k], G.beta);
        zd[k] = window_deriv(G.kernel, zu[k], G.beta) * zs[k];
        zo[k] = (G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[2]) : g) * G.cstride[2];
      }
    }
    float acc[3] = {0.f, 0.f, 0.f};
    for (int64_t b0 = 0; b0 < nbatch; b0 += NB) {
      const int nb = (int)(nbatch - b0 < NB ? nbatch - b0 : NB);
      typename V::T h[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < nb) h[j] = V::load(pvals, (b0 + j) * npts + p);
#pragma unroll 1
      for (int kx = 0; kx < r[0].count; ++kx) {
        float xu = 0.f, xs = 0.f, xw = 1.f, xd = 0.f;
        int64_t ox = 0;
        if (NDIM >= 3) {
          const int64_t g = r[0].g0 + kx;
          const float diff = t[0] - (float)g;
          xu = fminf(fabsf(diff) * G.inv_half[0], 1.0f);
          xs = (diff > 0.f ? 1.f : (diff < 0.f ? -1.f : 0.f)) * G.inv_half[0];
          xw = window(G.kernel, xu, G.beta);
          xd = window_deriv(G.kernel, xu, G.beta) * xs;
          ox = (G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[0]) : g) * G.cstride[0];
        }
#pragma unroll 1
        for (int ky = 0; ky < r[1].count; ++ky) {
          float yu = 0.f, ys = 0.f, yw = 1.f, yd = 0.f;
          int64_t oxy = ox;
          if (NDIM >= 2) {
            const int64_t g = r[1].g0 + ky;
            const float diff = t[1] - (float)g;
            yu = fminf(fabsf(diff) * G.inv_half[1], 1.0f);
            ys = (diff > 0.f ? 1.f : (diff < 0.f ? -1.f : 0.f)) * G.inv_half[1];
            yw = window(G.kernel, yu, G.beta);
            yd = window_deriv(G.kernel, yu, G.beta) * ys;
            oxy += (G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[1]) : g) * G.cstride[1];
          }
#pragma unroll
          for (int kz = 0; kz < GRAD_ZT; ++kz) {
            if (kz >= r[2].count) break;
            float s = 0.f;
            const int64_t off = oxy + zo[kz];
#pragma unroll
            for (int j = 0; j < NB; ++j)
              if (j < nb) s += V::rdot(h[j], V::load(gvals, (b0 + j) * G.ncells + off));
            if (G.norm == 1) {
              acc[0] += s * xd * yw * zw[kz];
              acc[1] += s * xw * yd * zw[kz];
              acc[2] += s * xw * yw * zd[kz];
            } else {
              const float rr = sqrtf(xu * xu + yu * yu + zu[kz] * zu[kz]);
              if (rr <= 1.0f && rr > 0.0f) {
                const float f = s * window_deriv(G.kernel, rr, G.beta) / rr;
                acc[0] += f * xu * xs;
                acc[1] += f * yu * ys;
                acc[2] += f * zu[kz] * zs[kz];
              }
            }
          }
        }
      }
    }
#pragma unroll
    for (int d = 3 - NDIM; d < 3; ++d) grad[p * NDIM + d - (3 - NDIM)] = acc[d];
  }
}

int interp_grad_locs(const float* gvals, const float* pvals, const float* locs,
                     float* grad, int64_t nbatch, int64_t npts,
                     const InterpGeom& G, int is_complex, hipStream_t s) {
  for (int d = 0; d < 3; ++d)
    if (G.ntap[d] > GRAD_ZT)
      return fail(TNL_EINVAL, "trajectory gradient supports kernel widths below 16");
  int64_t blocks = (npts + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
#define TNL_GRAD_LAUNCH(ND, CX)                                               \
  grad_locs_kernel<ND, CX><<<(unsigned)blocks, 256, 0, s>>>(gvals, pvals, locs, \
                                                           grad, nbatch, npts, G)
  if (G.ndim == 1) {
    if (is_complex) TNL_GRAD_LAUNCH(1, true); else TNL_GRAD_LAUNCH(1, false);
  } else if (G.ndim == 2) {
    if (is_complex) TNL_GRAD_LAUNCH(2, true); else TNL_GRAD_LAUNCH(2, false);
  } else {
    if (is_complex) TNL_GRAD_LAUNCH(3, true); else TNL_GRAD_LAUNCH(3, false);
  }
#undef TNL_GRAD_LAUNCH
  TNL_LAUNCH_CHECK("grad_locs");
  return TNL_OK;
}

}  // namespace tnl
