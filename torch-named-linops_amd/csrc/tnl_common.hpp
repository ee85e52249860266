// Shared device helpers and host error plumbing for libtnl.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <cstdio>
#include <string>

#include "tnl.h"

namespace tnl {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

#define TNL_LAUNCH_CHECK(what)                                     \
  do {                                                             \
    hipError_t _e = hipGetLastError();                             \
    if (_e != hipSuccess) return ::tnl::check_hip(_e, what);       \
  } while (0)

// ------------------------------------------------- streaming accesses
// Stores and loads marked non-temporal (the `nt` bit of the vector memory
// instruction): for data that streams through once, so it does not displace
// data that is reused from the caches.  Plain vector memory instructions;
// any type of 1, 2, 4, 8 or 16 bytes.  Measured on array_to_blocks (8 x
// 128^3 complex64, 8^3 blocks, stride 4: 976 MB of blocks written, the
// image re-read 8 times from L2): 303 -> 180 us with streaming stores.
template <typename T>
__device__ __forceinline__ void stream_store(T* p, const T& v) {
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  if constexpr (sizeof(T) == 16) {
    __builtin_nontemporal_store(*reinterpret_cast<const u4*>(&v), reinterpret_cast<u4*>(p));
  } else if constexpr (sizeof(T) == 8) {
    __builtin_nontemporal_store(*reinterpret_cast<const u2*>(&v), reinterpret_cast<u2*>(p));
  } else if constexpr (sizeof(T) == 4) {
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned*>(&v),
                                reinterpret_cast<unsigned*>(p));
  } else {
    *p = v;
  }
}
template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  T v;
  if constexpr (sizeof(T) == 16) {
    *reinterpret_cast<u4*>(&v) = __builtin_nontemporal_load(reinterpret_cast<const u4*>(p));
  } else if constexpr (sizeof(T) == 8) {
    *reinterpret_cast<u2*>(&v) = __builtin_nontemporal_load(reinterpret_cast<const u2*>(p));
  } else if constexpr (sizeof(T) == 4) {
    *reinterpret_cast<unsigned*>(&v) =
        __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(p));
  } else {
    v = *p;
  }
  return v;
}

// ------------------------------------------------------- kernel weights
// 16^k / (k!)^2: coefficients of I0(z) = sum_k u^k 16^k / (k!)^2 with
// u = z^2 / 64.  The factor 16^k keeps every coefficient up to k = 31 a
// normal float (1 / (k!)^2 alone underflows past k = 21).
constexpr double i0_coef(int k) {
  double f = 1.0;
  for (int i = 1; i <= k; ++i) f *= 16.0 / ((double)i * (double)i);
  return f;
}

template <int K>
__device__ __forceinline__ float i0_horner(float u, float p) {
  constexpr float c = (float)i0_coef(K);
  p = fmaf(p, u, c);
  if constexpr (K == 0) return p;
  else return i0_horner<K - 1>(u, p);
}

// sum_{k < N} u^k 16^k / (k!)^2, Horner in u (every term positive, so no
// cancellation).
template <int N>
__device__ __forceinline__ float i0_series(float u) {
  return i0_horner<N - 2>(u, (float)i0_coef(N - 1));
}

// Kaiser-Bessel window I0(beta * sqrt(1 - x^2)), unnormalised, |x| <= 1, as
// in src/torchlinops/functional/_interp/kernels.py:50-89.  For beta <= 30
// the power series of I0 in z^2 = beta^2 (1 - x^2) is summed directly: 16
// terms leave a tail below 1e-7 of I0(beta) for beta <= 10, 24 terms for
// beta <= 20, 32 terms for beta <= 30 (the first omitted term bounds the
// tail, which then decays by z^2 / (4 (k + 1)^2) per term).  That is
// branch-free and needs no sqrt or exp; in float32 it stays within 1e-6 of
// I0, as close as the Abramowitz & Stegun 9.8.1 / 9.8.2 polynomials
// (Handbook of Mathematical Functions, p. 378) that the reference
// evaluates and that serve larger beta here.
// Callers guarantee |x| <= 1 (clamped).
__device__ __forceinline__ float kb_window(float x, float beta) {
  const float v = fmaxf(1.0f - x * x, 0.0f);
  if (beta <= 10.0f) return i0_series<16>(beta * beta * (1.0f / 64.0f) * v);
  if (beta <= 20.0f) return i0_series<24>(beta * beta * (1.0f / 64.0f) * v);
  if (beta <= 30.0f) return i0_series<32>(beta * beta * (1.0f / 64.0f) * v);
  const float z = beta * sqrtf(v);
  if (z < 3.75f) {
    const float t = z * (1.0f / 3.75f);
    const float q = t * t;
    return 1.0f +
           q * (3.5156229f +
                q * (3.0899424f +
                     q * (1.2067492f +
                          q * (0.2659732f + q * (0.0360768f + q * 0.0045813f)))));
  }
  // z >= 3.75 here, so the reciprocal and rsqrt are of well-scaled values
  const float rz = __frcp_rn(z);
  const float r = 3.75f * rz;  // 1/t
  const float poly =
      0.39894228f +
      r * (0.01328592f +
           r * (0.00225319f +
                r * (-0.00157565f +
                     r * (0.00916281f +
                          r * (-0.02057706f +
                               r * (0.02635537f +
                                    r * (-0.01647633f + r * 0.00392377f)))))));
  return expf(z) * sqrtf(rz) * poly;
}

// Linear-interpolation hat max(1 - |x|, 0) (kernels.py:92-94).
__device__ __forceinline__ float spline_window(float x) {
  return fmaxf(1.0f - fabsf(x), 0.0f);
}

__device__ __forceinline__ float window(int kernel, float x, float beta) {
  return kernel == TNL_KERNEL_KAISER_BESSEL ? kb_window(x, beta)
                                            : spline_window(x);
}

// d/du of the Kaiser-Bessel window I0(beta * sqrt(1 - u^2)) for u in [0, 1]:
//   -beta^2 * u * I1(z) / z,  z = beta * sqrt(1 - u^2),
// which stays finite at u = 1 (I1(z)/z -> 1/2).  I1(z)/z is the Abramowitz &
// Stegun 9.8.3 polynomial for z < 3.75 and 9.8.4 (z^1/2 e^-z I1(z)) above,
// the companions of the I0 approximations used by kb_window.
__device__ __forceinline__ float kb_window_deriv(float u, float beta) {
  const float z = beta * sqrtf(fmaxf(1.0f - u * u, 0.0f));
  float g;  // I1(z) / z
  if (z < 3.75f) {
    const float t = z * (1.0f / 3.75f);
    const float q = t * t;
    g = 0.5f +
        q * (0.87890594f +
             q * (0.51498869f +
                  q * (0.15084934f +
                       q * (0.02658733f + q * (0.00301532f + q * 0.00032411f)))));
  } else {
    const float rz = __frcp_rn(z);
    const float r = 3.75f * rz;
    const float poly =
        0.39894228f +
        r * (-0.03988024f +
             r * (-0.00362018f +
                  r * (0.00163801f +
                       r * (-0.01031555f +
                            r * (0.02282967f +
                                 r * (-0.02895312f +
                                      r * (0.01787654f + r * -0.00420059f)))))));
    g = expf(z) * sqrtf(rz) * rz * poly;
  }
  return -beta * beta * u * g;
}

// d/du of the hat max(1 - u, 0) on its support u in [0, 1].
__device__ __forceinline__ float spline_window_deriv(float u) {
  return u <= 1.0f ? -1.0f : 0.0f;
}

__device__ __forceinline__ float window_deriv(int kernel, float u, float beta) {
  return kernel == TNL_KERNEL_KAISER_BESSEL ? kb_window_deriv(u, beta)
                                            : spline_window_deriv(u);
}

__device__ __forceinline__ int64_t mod_pos(int64_t g, int64_t n) {
  int64_t r = g % n;
  return r < 0 ? r + n : r;
}

// Per-launch geometry of one interpolation problem.
struct InterpGeom {
  int64_t n[3];        // grid size per dim (dims padded to 3 on the left)
  int64_t cstride[3];  // cell stride per dim, in cells
  int64_t ncells;      // prod(n)
  float half[3];       // w_d / 2
  float inv_half[3];   // 2 / w_d
  int ntap[3];         // ceil(w_d) + 1
  int ndim;
  int kernel;
  int norm;
  int pad;
  float beta;
};

// One dimension of the tap set of a point at coordinate t:
// g = ceil(t - w/2) + k for k in [0, ntap); a tap is kept iff
// |g - t| <= w/2 (kernels.py:264-279, 383-388); under zero padding taps
// outside [0, n) are dropped, under circular padding g wraps modulo n.
// Returns the index of the first tap, the number of consecutive kept taps
// and, for k < count, the distance (normalised by w/2) of each kept tap.
struct TapRange {
  int64_t g0;  // first kept tap, unwrapped
  int count;
};

__device__ __forceinline__ TapRange tap_range(float t, float half, int ntap,
                                              int64_t n, int pad) {
  TapRange r;
  r.g0 = 0;
  r.count = 0;
  // Non-finite coordinates have no taps (every |g - t| <= w/2 test fails).
  if (!(fabsf(t) < 4.0e18f)) return r;
  const int64_t lower = (int64_t)ceilf(t - half);
  // kept taps form one interval; find it by testing the ends
  int k0 = 0, k1 = ntap - 1;
  while (k0 <= k1 && !(fabsf((float)(lower + k0) - t) <= half)) ++k0;
  while (k1 >= k0 && !(fabsf((float)(lower + k1) - t) <= half)) --k1;
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

// Cells [0, tl) (tl <= 32) of a tile starting at cell `origin` of an axis
// of n cells that the taps of coordinate t fall on (bit l = cell origin +
// l; taps wrap under circular padding).  The tile-local footprint of one
// axis.
__device__ __forceinline__ unsigned tile_cells(float t, float half, int ntap,
                                               int64_t n, int pad, int origin, int tl) {
  const TapRange r = tap_range(t, half, ntap, n, pad);
  int64_t a = r.g0;
  if (pad == TNL_PAD_CIRCULAR) {
    if (a < 0) a += n;
    else if (a >= n) a -= n;
    if (a < 0 || a >= n) a = mod_pos(r.g0, n);
  }
  const int lo = (int)a - origin;
  unsigned m = 0;
  for (int k = 0; k < r.count; ++k) {
    int l = lo + k;
    if (pad == TNL_PAD_CIRCULAR && origin + l >= n) l -= (int)n;
    if ((unsigned)l < (unsigned)tl) m |= 1u << l;
  }
  return m;
}

// Blocks of a 16x32 (y, z) tile that a footprint with tile-local cell masks
// (my: 16 cells, mz: 32 cells) touches: bit 4 u + v for the 4x4 block of
// cells 4 u .. 4 u + 3 along z and 4 v .. 4 v + 3 along y.
__device__ __forceinline__ unsigned tile_blocks32_2d(unsigned my, unsigned mz) {
  unsigned yb = 0;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    if ((my >> (4 * v)) & 15u) yb |= 1u << v;
  unsigned m = 0;
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if ((mz >> (4 * u)) & 15u) m |= yb << (4 * u);
  return m;
}

// Blocks of an 8x8x8 tile that a footprint with tile-local cell masks
// (mx, my, mz) touches: bit 4 x + 2 py + pz for x plane x and the 4x4
// (y, z) patch (py, pz) (cells 4 py .. 4 py + 3 along y, likewise z).
__device__ __forceinline__ unsigned tile_blocks32(unsigned mx, unsigned my, unsigned mz) {
  const unsigned yh = ((my & 15u) ? 1u : 0u) | ((my >> 4) ? 2u : 0u);
  const unsigned zh = ((mz & 15u) ? 1u : 0u) | ((mz >> 4) ? 2u : 0u);
  const unsigned pm = ((yh & 1u) && (zh & 1u) ? 1u : 0u) | ((yh & 1u) && (zh & 2u) ? 2u : 0u) |
                      ((yh & 2u) && (zh & 1u) ? 4u : 0u) | ((yh & 2u) && (zh & 2u) ? 8u : 0u);
  unsigned m = 0;
#pragma unroll
  for (int x = 0; x < 8; ++x)
    if ((mx >> x) & 1u) m |= pm << (4 * x);
  return m;
}

}  // namespace tnl
