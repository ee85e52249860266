// Nonuniform interpolation (ungrid) and gridding for gfx950.
//
// Two code paths:
//   * direct  -- one thread per point, weights in registers, gathers from
//               (ungrid) / fp32 atomics into (grid) HBM.  Used without a
//               plan, for 1-D problems and for kernel widths > 7.
//   * binned  -- points pre-sorted into spatial tiles by a plan
//               (tnl_plan.hip); see tnl_binned.hip.
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

// Distance |g - t| / (w/2) of tap g (clamped to 1 so KB never sees |x|>1)
// and its cell offset along padded axis d.
__device__ __forceinline__ float tap_dist(int64_t g, float t,
                                          const InterpGeom& G, int d) {
  return fminf(fabsf((float)g - t) * G.inv_half[d], 1.0f);
}
__device__ __forceinline__ int64_t tap_off(int64_t g, const InterpGeom& G,
                                           int d) {
  const int64_t gw = G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[d]) : g;
  return gw * G.cstride[d];
}

template <bool CPLX>
struct Val;
template <>
struct Val<false> {
  using T = float;
  __device__ static T zero() { return 0.0f; }
  __device__ static T load(const float* p, int64_t i) { return p[i]; }
  __device__ static void store(float* p, int64_t i, T v) { p[i] = v; }
  __device__ static void fma(T& acc, float w, T v) { acc = fmaf(w, v, acc); }
  __device__ static void atomic_add(float* p, int64_t i, float w, T v) {
    unsafeAtomicAdd(p + i, w * v);
  }
};
template <>
struct Val<true> {
  using T = float2;
  __device__ static T zero() { return make_float2(0.0f, 0.0f); }
  __device__ static T load(const float* p, int64_t i) {
    return reinterpret_cast<const float2*>(p)[i];
  }
  __device__ static void store(float* p, int64_t i, T v) {
    reinterpret_cast<float2*>(p)[i] = v;
  }
  __device__ static void fma(T& acc, float w, T v) {
    acc.x = fmaf(w, v.x, acc.x);
    acc.y = fmaf(w, v.y, acc.y);
  }
  __device__ static void atomic_add(float* p, int64_t i, float w, T v) {
    unsafeAtomicAdd(p + 2 * i, w * v.x);
    unsafeAtomicAdd(p + 2 * i + 1, w * v.y);
  }
};

// One thread per point.  Batches are processed NB at a time so the tap
// weights are computed once per NB channels.  The innermost (z) axis keeps
// its weights in registers when ZT (its compile-time tap capacity) is small;
// the outer axes are walked with runtime loops.  ADJ selects gridding
// (atomic scatter) instead of interpolation (gather).
template <int NDIM, int ZT, bool CPLX, bool ADJ>
__global__ __launch_bounds__(256) void interp_direct_kernel(
    const float* __restrict__ vals, const float* __restrict__ locs,
    float* __restrict__ out, int64_t nbatch, int64_t npts, InterpGeom G) {
  using V = Val<CPLX>;
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
    // z-axis cache: window value (norm 1) or distance (norm 2), offset
    float zw[ZT];
    int64_t zo[ZT];
#pragma unroll
    for (int k = 0; k < ZT; ++k) {
      zw[k] = 0.f;
      zo[k] = 0;
      if (k < r[2].count) {
        const int64_t g = r[2].g0 + k;
        const float x = tap_dist(g, t[2], G, 2);
        zw[k] = G.norm == 1 ? window(G.kernel, x, G.beta) : x;
        zo[k] = tap_off(g, G, 2);
      }
    }
    for (int64_t b0 = 0; b0 < nbatch; b0 += NB) {
      const int nb = (int)(nbatch - b0 < NB ? nbatch - b0 : NB);
      typename V::T acc[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        acc[j] = V::zero();
        if (ADJ && j < nb) acc[j] = V::load(vals, (b0 + j) * npts + p);
      }
#pragma unroll 1
      for (int kx = 0; kx < r[0].count; ++kx) {
        float ax = 1.0f, dx = 0.0f;
        int64_t ox = 0;
        if (NDIM >= 3) {
          const int64_t g = r[0].g0 + kx;
          dx = tap_dist(g, t[0], G, 0);
          ax = G.norm == 1 ? window(G.kernel, dx, G.beta) : 1.0f;
          ox = tap_off(g, G, 0);
        }
#pragma unroll 1
        for (int ky = 0; ky < r[1].count; ++ky) {
          float axy = ax, dy = 0.0f;
          int64_t oxy = ox;
          if (NDIM >= 2) {
            const int64_t g = r[1].g0 + ky;
            dy = tap_dist(g, t[1], G, 1);
            if (G.norm == 1) axy *= window(G.kernel, dy, G.beta);
            oxy += tap_off(g, G, 1);
          }
#pragma unroll
          for (int kz = 0; kz < ZT; ++kz) {
            if (kz >= r[2].count) break;
            float w;
            if (G.norm == 1) {
              w = axy * zw[kz];
            } else {
              const float rr = sqrtf(dx * dx + dy * dy + zw[kz] * zw[kz]);
              w = rr <= 1.0f ? window(G.kernel, rr, G.beta) : 0.0f;
            }
            const int64_t off = oxy + zo[kz];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
              if (j >= nb) break;
              if (ADJ)
                V::atomic_add(out, (b0 + j) * G.ncells + off, w, acc[j]);
              else
                V::fma(acc[j], w, V::load(vals, (b0 + j) * G.ncells + off));
            }
          }
        }
      }
      if (!ADJ) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
          if (j < nb) V::store(out, (b0 + j) * npts + p, acc[j]);
      }
    }
  }
}

// ---- float64 / complex128 interpolation (no plan; one thread per point)
// Windows in double: Kaiser-Bessel I0 from its power series
// sum_k (z^2 / 4)^k / (k!)^2, summed until a term no longer changes the
// double sum (the terms fall monotonically once k > z / 2).
__device__ __forceinline__ double kb_window_f64(double x, double beta) {
  const double v = fmax(1.0 - x * x, 0.0);
  const double q = 0.25 * beta * beta * v;  // z^2 / 4
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 400; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-17 * sum && k > 2) break;
  }
  return sum;
}

__device__ __forceinline__ double window_f64(int kernel, double x, double beta) {
  return kernel == TNL_KERNEL_KAISER_BESSEL ? kb_window_f64(x, beta) : fmax(1.0 - fabs(x), 0.0);
}

// Taps of coordinate t along one padded axis in double: lower =
// ceil(t - w/2), kept iff |t - g| <= w/2, zero padding drops cells outside
// [0, n); non-finite coordinates have none.
struct TapRangeD {
  int64_t g0;
  int count;
};

__device__ __forceinline__ TapRangeD tap_range_f64(double t, double half, int ntap, int64_t n,
                                                   int pad) {
  TapRangeD r{0, 0};
  if (!(fabs(t) < 4.0e18)) return r;
  const int64_t lower = (int64_t)ceil(t - half);
  int k0 = 0, k1 = ntap - 1;
  while (k0 <= k1 && !(fabs((double)(lower + k0) - t) <= half)) ++k0;
  while (k1 >= k0 && !(fabs((double)(lower + k1) - t) <= half)) --k1;
  if (k1 < k0) return r;
  int64_t a = lower + k0, b = lower + k1;
  if (pad == TNL_PAD_ZERO) {
    a = a < 0 ? 0 : a;
    b = b > n - 1 ? n - 1 : b;
    if (b < a) return r;
  }
  r.g0 = a;
  r.count = (int)(b - a + 1);
  return r;
}

template <int NDIM, bool CPLX, bool ADJ>
__global__ __launch_bounds__(256) void interp_direct_f64_kernel(
    const double* __restrict__ vals, const double* __restrict__ locs, double* __restrict__ out,
    int64_t nbatch, int64_t npts, InterpGeom G, double beta) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npts; p += stride) {
    double t[3] = {0.0, 0.0, 0.0}, half[3], invh[3];
    TapRangeD r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      half[d] = (double)G.half[d];
      invh[d] = 1.0 / half[d];
      r[d].g0 = 0;
      r[d].count = 1;
      if (d >= 3 - NDIM) {
        t[d] = locs[p * NDIM + d - (3 - NDIM)];
        r[d] = tap_range_f64(t[d], half[d], G.ntap[d], G.n[d], G.pad);
      }
    }
    for (int64_t b = 0; b < nbatch; ++b) {
      double ar = 0.0, ai = 0.0;
      if (ADJ) {
        ar = CPLX ? vals[2 * (b * npts + p)] : vals[b * npts + p];
        ai = CPLX ? vals[2 * (b * npts + p) + 1] : 0.0;
      }
      double sr = 0.0, si = 0.0;
      for (int kx = 0; kx < r[0].count; ++kx) {
        double dx = 0.0, wx = 1.0;
        int64_t ox = 0;
        if (NDIM >= 3) {
          const int64_t g = r[0].g0 + kx;
          dx = fmin(fabs((double)g - t[0]) * invh[0], 1.0);
          wx = G.norm == 1 ? window_f64(G.kernel, dx, beta) : 1.0;
          ox = (G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[0]) : g) * G.cstride[0];
        }
        for (int ky = 0; ky < r[1].count; ++ky) {
          double dy = 0.0, wxy = wx;
          int64_t oxy = ox;
          if (NDIM >= 2) {
            const int64_t g = r[1].g0 + ky;
            dy = fmin(fabs((double)g - t[1]) * invh[1], 1.0);
            if (G.norm == 1) wxy *= window_f64(G.kernel, dy, beta);
            oxy += (G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[1]) : g) * G.cstride[1];
          }
          for (int kz = 0; kz < r[2].count; ++kz) {
            const int64_t g = r[2].g0 + kz;
            const double dz = fmin(fabs((double)g - t[2]) * invh[2], 1.0);
            double w;
            if (G.norm == 1) {
              w = wxy * window_f64(G.kernel, dz, beta);
            } else {
              const double rr = sqrt(dx * dx + dy * dy + dz * dz);
              w = rr <= 1.0 ? window_f64(G.kernel, rr, beta) : 0.0;
            }
            if (!isfinite(w)) w = 0.0;
            const int64_t off =
                oxy + (G.pad == TNL_PAD_CIRCULAR ? mod_pos(g, G.n[2]) : g) + b * G.ncells;
            if (ADJ) {
              if (CPLX) {
                unsafeAtomicAdd(out + 2 * off, w * ar);
                unsafeAtomicAdd(out + 2 * off + 1, w * ai);
              } else {
                unsafeAtomicAdd(out + off, w * ar);
              }
            } else {
              if (CPLX) {
                sr = fma(w, vals[2 * off], sr);
                si = fma(w, vals[2 * off + 1], si);
              } else {
                sr = fma(w, vals[off], sr);
              }
            }
          }
        }
      }
      if (!ADJ) {
        if (CPLX) {
          out[2 * (b * npts + p)] = sr;
          out[2 * (b * npts + p) + 1] = si;
        } else {
          out[b * npts + p] = sr;
        }
      }
    }
  }
}

int interp_direct_f64(bool adjoint, const double* vals, const double* locs, double* out,
                      int64_t nbatch, int64_t npts, const InterpGeom& G, int is_complex,
                      double beta, hipStream_t s) {
  if (npts == 0 || nbatch == 0) return TNL_OK;
  int64_t nb = (npts + 255) / 256;
  if (nb > 65536) nb = 65536;
#define TNL_F64_LAUNCH(ND)                                                                 \
  {                                                                                        \
    if (adjoint) {                                                                         \
      if (is_complex)                                                                      \
        interp_direct_f64_kernel<ND, true, true><<<(unsigned)nb, 256, 0, s>>>(             \
            vals, locs, out, nbatch, npts, G, beta);                                       \
      else                                                                                 \
        interp_direct_f64_kernel<ND, false, true><<<(unsigned)nb, 256, 0, s>>>(            \
            vals, locs, out, nbatch, npts, G, beta);                                       \
    } else {                                                                               \
      if (is_complex)                                                                      \
        interp_direct_f64_kernel<ND, true, false><<<(unsigned)nb, 256, 0, s>>>(            \
            vals, locs, out, nbatch, npts, G, beta);                                       \
      else                                                                                 \
        interp_direct_f64_kernel<ND, false, false><<<(unsigned)nb, 256, 0, s>>>(           \
            vals, locs, out, nbatch, npts, G, beta);                                       \
    }                                                                                      \
    TNL_LAUNCH_CHECK("interp_direct_f64");                                                 \
    return TNL_OK;                                                                         \
  }
  switch (G.ndim) {
    case 1: TNL_F64_LAUNCH(1)
    case 2: TNL_F64_LAUNCH(2)
    case 3: TNL_F64_LAUNCH(3)
  }
#undef TNL_F64_LAUNCH
  return fail(TNL_EINVAL, "ndim must be 1, 2 or 3");
}

__global__ void fill_zero_kernel(float4* __restrict__ p, int64_t n4,
                                 float* __restrict__ tail, int ntail) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += stride)
    p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0.f;
}

int fill_zero(float* p, int64_t n, hipStream_t s) {
  if (n <= 0) return TNL_OK;
  // vector stores from the first 16-byte boundary; the head and tail
  // (fewer than 4 floats each) are written by the first block
  int head = (int)(((16 - ((uintptr_t)p & 15)) & 15) / 4);
  if ((uintptr_t)p & 3) return fail(TNL_EINVAL, "fill_zero: unaligned float pointer");
  if (head > n) head = (int)n;
  if (head > 0) {
    fill_zero_kernel<<<1, 64, 0, s>>>(nullptr, 0, p, head);
    TNL_LAUNCH_CHECK("fill_zero");
    p += head;
    n -= head;
  }
  const int64_t n4 = n / 4;
  const int ntail = (int)(n - n4 * 4);
  int64_t blocks = (n4 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  fill_zero_kernel<<<(unsigned)blocks, 256, 0, s>>>(
      reinterpret_cast<float4*>(p), n4, p + n4 * 4, ntail);
  TNL_LAUNCH_CHECK("fill_zero");
  return TNL_OK;
}

static unsigned direct_blocks(int64_t npts) {
  int64_t b = (npts + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

template <int NDIM, int ZT>
static int launch_direct(bool adjoint, const float* vals, const float* locs,
                         float* out, int64_t nbatch, int64_t npts,
                         const InterpGeom& G, int is_complex, hipStream_t s) {
  const unsigned nb = direct_blocks(npts);
  if (!adjoint) {
    if (is_complex)
      interp_direct_kernel<NDIM, ZT, true, false>
          <<<nb, 256, 0, s>>>(vals, locs, out, nbatch, npts, G);
    else
      interp_direct_kernel<NDIM, ZT, false, false>
          <<<nb, 256, 0, s>>>(vals, locs, out, nbatch, npts, G);
    TNL_LAUNCH_CHECK("ungrid_direct");
  } else {
    if (is_complex)
      interp_direct_kernel<NDIM, ZT, true, true>
          <<<nb, 256, 0, s>>>(vals, locs, out, nbatch, npts, G);
    else
      interp_direct_kernel<NDIM, ZT, false, true>
          <<<nb, 256, 0, s>>>(vals, locs, out, nbatch, npts, G);
    TNL_LAUNCH_CHECK("grid_direct");
  }
  return TNL_OK;
}

int interp_direct(bool adjoint, const float* vals, const float* locs,
                  float* out, int64_t nbatch, int64_t npts,
                  const InterpGeom& G, int is_complex, hipStream_t s) {
  if (npts == 0 || nbatch == 0) return TNL_OK;
  // z (innermost) taps cached in registers: capacity 8 or 64
#define TNL_DISPATCH_T(ND)                                                  \
  if (G.ntap[2] <= 8)                                                       \
    return launch_direct<ND, 8>(adjoint, vals, locs, out, nbatch, npts, G,  \
                                is_complex, s);                             \
  else                                                                      \
    return launch_direct<ND, TNL_MAX_TAPS>(adjoint, vals, locs, out,        \
                                           nbatch, npts, G, is_complex, s);
  switch (G.ndim) {
    case 1: TNL_DISPATCH_T(1)
    case 2: TNL_DISPATCH_T(2)
    case 3: TNL_DISPATCH_T(3)
  }
#undef TNL_DISPATCH_T
  return fail(TNL_EINVAL, "ndim must be 1, 2 or 3");
}

}  // namespace tnl
