// Centred zero-pad fused with ifftshift, and fftshift fused with the
// centred crop, for gfx950: the two data-movement passes around the FFTs
// of a Toeplitz normal operator.
//
// Image index i_d in [0, N_d) sits at centred coordinate c_d = i_d - N_d/2
// (integer division).  The centred pad puts it at c_d + P_d/2 of the
// padded grid and ifftshift moves that to q_d = c_d mod P_d, so
//   pad_wrapped:  y[n, q] = x[n, i] if every axis has an image index whose
//                 c_d == q_d (mod P_d), else 0;
//   crop_wrapped: x[n, i] = y[n, (i - N/2) mod P]  (fftshift then crop).
// An optional real weight w[i] (float32, one per image cell, shared by
// the batch) multiplies each copied value: the NUFFT's apodization and
// 1/scale folded into the same pass (float32 and complex64 data).
// Both are HBM-bound copies.  A workgroup walks whole rows of the last
// axis: the row's outer coordinates are decoded once, lanes move
// consecutive elements, so reads and writes are coalesced except at the
// single wrap point of a row.
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

namespace {

template <int EB>
struct WElem;
template <>
struct WElem<4> { using T = uint32_t; };
template <>
struct WElem<8> { using T = uint2; };
template <>
struct WElem<16> { using T = uint4; };

struct WrapGeom {
  int64_t N[3], P[3];  // left-padded with 1
  int64_t nbatch;
};

// image index of padded-grid coordinate q on an axis, or -1
__device__ __forceinline__ int64_t wrap_source(int64_t q, int64_t N, int64_t P) {
  int64_t c = q;
  if (c + N / 2 >= N) c -= P;
  const int64_t i = c + N / 2;
  return (i >= 0 && i < N) ? i : -1;
}

// value times a real weight: float32 and interleaved complex64 elements
__device__ __forceinline__ uint32_t scaled(uint32_t v, float w) {
  return __float_as_uint(__uint_as_float(v) * w);
}
__device__ __forceinline__ uint2 scaled(uint2 v, float w) {
  return make_uint2(__float_as_uint(__uint_as_float(v.x) * w),
                    __float_as_uint(__uint_as_float(v.y) * w));
}
__device__ __forceinline__ uint4 scaled(uint4 v, float) { return v; }  // not launched

// I: type of the row number and its div / mod (uint32_t when every row
// number fits, avoiding the 64-bit division sequence per row)
template <typename E, bool W, typename I>
__global__ __launch_bounds__(256) void pad_wrapped_kernel(const E* __restrict__ x,
                                                          E* __restrict__ y,
                                                          const float* __restrict__ w,
                                                          WrapGeom g) {
  const I rows = (I)(g.nbatch * g.P[0] * g.P[1]);
  const I p0 = (I)g.P[0], p1 = (I)g.P[1];
  for (I r = blockIdx.x; r < rows; r += gridDim.x) {
    const int64_t q1 = (int64_t)(r % p1);
    const I t = r / p1;
    const int64_t q0 = (int64_t)(t % p0);
    const int64_t n = (int64_t)(t / p0);
    const int64_t i0 = wrap_source(q0, g.N[0], g.P[0]);
    const int64_t i1 = wrap_source(q1, g.N[1], g.P[1]);
    E* dst = y + (int64_t)r * g.P[2];
    const bool row_ok = i0 >= 0 && i1 >= 0;
    const int64_t img_row = ((row_ok ? i0 : 0) * g.N[1] + (row_ok ? i1 : 0)) * g.N[2];
    const E* src = x + n * g.N[0] * g.N[1] * g.N[2] + img_row;
    for (int64_t q2 = threadIdx.x; q2 < g.P[2]; q2 += blockDim.x) {
      E v{};
      if (row_ok) {
        const int64_t i2 = wrap_source(q2, g.N[2], g.P[2]);
        if (i2 >= 0) {
          v = src[i2];
          if (W) v = scaled(v, w[img_row + i2]);
        }
      }
      dst[q2] = v;
    }
  }
}

// complex64 pad with 16-byte stores: a lane writes the cell pair
// (2 j, 2 j + 1) of a row (P2 even); a workgroup pass covers
// max(1, 256 / (P2 / 2)) whole rows, or one row in 256-pair steps.
template <bool W, typename I>
__global__ __launch_bounds__(256) void pad_wrapped_pairs_kernel(const uint2* __restrict__ x,
                                                                uint4* __restrict__ y,
                                                                const float* __restrict__ w,
                                                                WrapGeom g) {
  const I rows = (I)(g.nbatch * g.P[0] * g.P[1]);
  const I p0 = (I)g.P[0], p1 = (I)g.P[1];
  const int pairs = (int)(g.P[2] / 2);
  const int per = pairs >= 256 ? 1 : 256 / pairs;  // rows per pass
  const int lr = (int)threadIdx.x / (pairs >= 256 ? 256 : pairs);
  const int lp = (int)threadIdx.x % (pairs >= 256 ? 256 : pairs);
  if (lr >= per) return;  // lanes past the last whole row of a pass
  for (I r = (I)blockIdx.x * per + lr; r < rows; r += (I)gridDim.x * per) {
    const int64_t q1 = (int64_t)(r % p1);
    const I t = r / p1;
    const int64_t q0 = (int64_t)(t % p0);
    const int64_t n = (int64_t)(t / p0);
    const int64_t i0 = wrap_source(q0, g.N[0], g.P[0]);
    const int64_t i1 = wrap_source(q1, g.N[1], g.P[1]);
    const bool row_ok = i0 >= 0 && i1 >= 0;
    const int64_t img_row = ((row_ok ? i0 : 0) * g.N[1] + (row_ok ? i1 : 0)) * g.N[2];
    const uint2* src = x + n * g.N[0] * g.N[1] * g.N[2] + img_row;
    uint4* dst = y + (int64_t)r * pairs;
    for (int j = lp; j < pairs; j += 256) {
      uint2 v[2] = {make_uint2(0u, 0u), make_uint2(0u, 0u)};
      if (row_ok) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t i2 = wrap_source(2 * j + e, g.N[2], g.P[2]);
          if (i2 >= 0) {
            v[e] = src[i2];
            if (W) v[e] = scaled(v[e], w[img_row + i2]);
          }
        }
      }
      dst[j] = make_uint4(v[0].x, v[0].y, v[1].x, v[1].y);
    }
  }
}

template <typename E, bool W, typename I>
__global__ __launch_bounds__(256) void crop_wrapped_kernel(const E* __restrict__ y,
                                                           E* __restrict__ x,
                                                           const float* __restrict__ w,
                                                           WrapGeom g) {
  const I rows = (I)(g.nbatch * g.N[0] * g.N[1]);
  const I n0 = (I)g.N[0], n1 = (I)g.N[1];
  for (I r = blockIdx.x; r < rows; r += gridDim.x) {
    const int64_t i1 = (int64_t)(r % n1);
    const I t = r / n1;
    const int64_t i0 = (int64_t)(t % n0);
    const int64_t n = (int64_t)(t / n0);
    // i - N/2 lies in [-N/2, N/2) and P >= N: one conditional wrap
    int64_t q0 = i0 - g.N[0] / 2, q1 = i1 - g.N[1] / 2;
    if (q0 < 0) q0 += g.P[0];
    if (q1 < 0) q1 += g.P[1];
    const E* src = y + ((n * g.P[0] + q0) * g.P[1] + q1) * g.P[2];
    E* dst = x + (int64_t)r * g.N[2];
    const float* wrow = W ? w + (i0 * g.N[1] + i1) * g.N[2] : nullptr;
    for (int64_t i2 = threadIdx.x; i2 < g.N[2]; i2 += blockDim.x) {
      int64_t q2 = i2 - g.N[2] / 2;
      if (q2 < 0) q2 += g.P[2];
      E v = src[q2];
      if (W) v = scaled(v, wrow[i2]);
      dst[i2] = v;
    }
  }
}

// complex64 crop with 16-byte stores: a lane writes the cell pair
// (2 j, 2 j + 1) of an output row (N2 even); the two sources are
// (2 j - N2/2) mod P2 and the next cell, adjacent unless the pair straddles
// the wrap.  A workgroup pass covers max(1, 256 / (N2 / 2)) rows.
template <bool W, typename I>
__global__ __launch_bounds__(256) void crop_wrapped_pairs_kernel(const uint2* __restrict__ y,
                                                                 uint4* __restrict__ x,
                                                                 const float* __restrict__ w,
                                                                 WrapGeom g) {
  const I rows = (I)(g.nbatch * g.N[0] * g.N[1]);
  const I n0 = (I)g.N[0], n1 = (I)g.N[1];
  const int pairs = (int)(g.N[2] / 2);
  const int per = pairs >= 256 ? 1 : 256 / pairs;
  const int lr = (int)threadIdx.x / (pairs >= 256 ? 256 : pairs);
  const int lp = (int)threadIdx.x % (pairs >= 256 ? 256 : pairs);
  if (lr >= per) return;
  for (I r = (I)blockIdx.x * per + lr; r < rows; r += (I)gridDim.x * per) {
    const int64_t i1 = (int64_t)(r % n1);
    const I t = r / n1;
    const int64_t i0 = (int64_t)(t % n0);
    const int64_t n = (int64_t)(t / n0);
    int64_t q0 = i0 - g.N[0] / 2, q1 = i1 - g.N[1] / 2;
    if (q0 < 0) q0 += g.P[0];
    if (q1 < 0) q1 += g.P[1];
    const uint2* src = y + ((n * g.P[0] + q0) * g.P[1] + q1) * g.P[2];
    const float* wrow = W ? w + (i0 * g.N[1] + i1) * g.N[2] : nullptr;
    uint4* dst = x + (int64_t)r * pairs;
    for (int j = lp; j < pairs; j += 256) {
      uint2 v[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        int64_t q2 = 2 * j + e - g.N[2] / 2;
        if (q2 < 0) q2 += g.P[2];
        v[e] = src[q2];
        if (W) v[e] = scaled(v[e], wrow[2 * j + e]);
      }
      dst[j] = make_uint4(v[0].x, v[0].y, v[1].x, v[1].y);
    }
  }
}

unsigned row_blocks(int64_t rows) {
  const int64_t cap = 256 * 64;
  return (unsigned)(rows < 1 ? 1 : (rows > cap ? cap : rows));
}

template <int EB, bool W>
int launch_wrap(bool pad, const void* in, void* out, const float* w, const WrapGeom& g,
                hipStream_t s) {
  using E = typename WElem<EB>::T;
  const int64_t rows = pad ? g.nbatch * g.P[0] * g.P[1] : g.nbatch * g.N[0] * g.N[1];
  // 32-bit row numbers when a grid-stride step (at most 16384 workgroups
  // x 256 rows) past the last row still fits
  const bool small = rows + 256LL * 64 * 256 < (1LL << 32);
  if (pad && EB == 8 && g.P[2] % 2 == 0 && ((uintptr_t)out & 15) == 0) {
    const int64_t pairs = g.P[2] / 2;
    const int64_t per = pairs >= 256 ? 1 : 256 / pairs;
    const unsigned nb = row_blocks((rows + per - 1) / per);
    if (small)
      pad_wrapped_pairs_kernel<W, uint32_t><<<nb, 256, 0, s>>>((const uint2*)in, (uint4*)out,
                                                              w, g);
    else
      pad_wrapped_pairs_kernel<W, int64_t><<<nb, 256, 0, s>>>((const uint2*)in, (uint4*)out,
                                                             w, g);
    TNL_LAUNCH_CHECK("pad_wrapped");
    return TNL_OK;
  }
  if (!pad && EB == 8 && g.N[2] % 2 == 0 && ((uintptr_t)out & 15) == 0) {
    const int64_t pairs = g.N[2] / 2;
    const int64_t per = pairs >= 256 ? 1 : 256 / pairs;
    const unsigned nb = row_blocks((rows + per - 1) / per);
    if (small)
      crop_wrapped_pairs_kernel<W, uint32_t><<<nb, 256, 0, s>>>((const uint2*)in, (uint4*)out,
                                                               w, g);
    else
      crop_wrapped_pairs_kernel<W, int64_t><<<nb, 256, 0, s>>>((const uint2*)in, (uint4*)out,
                                                              w, g);
    TNL_LAUNCH_CHECK("crop_wrapped");
    return TNL_OK;
  }
  if (pad) {
    if (small)
      pad_wrapped_kernel<E, W, uint32_t><<<row_blocks(rows), 256, 0, s>>>(
          (const E*)in, (E*)out, w, g);
    else
      pad_wrapped_kernel<E, W, int64_t><<<row_blocks(rows), 256, 0, s>>>(
          (const E*)in, (E*)out, w, g);
    TNL_LAUNCH_CHECK("pad_wrapped");
  } else {
    if (small)
      crop_wrapped_kernel<E, W, uint32_t><<<row_blocks(rows), 256, 0, s>>>(
          (const E*)in, (E*)out, w, g);
    else
      crop_wrapped_kernel<E, W, int64_t><<<row_blocks(rows), 256, 0, s>>>(
          (const E*)in, (E*)out, w, g);
    TNL_LAUNCH_CHECK("crop_wrapped");
  }
  return TNL_OK;
}

int wrap_entry(bool pad, const void* in, void* out, int elem_bytes, int64_t nbatch,
               int ndim, const int64_t* im_size, const int64_t* pad_size,
               const float* weight, void* stream) {
  if (ndim < 1 || ndim > 3) return fail(TNL_EINVAL, "ndim must be 1, 2 or 3");
  if (!im_size || !pad_size) return fail(TNL_EINVAL, "null size array");
  if (nbatch < 0) return fail(TNL_EINVAL, "nbatch must be >= 0");
  WrapGeom g;
  g.nbatch = nbatch;
  for (int d = 0; d < 3; ++d) g.N[d] = g.P[d] = 1;
  for (int d = 0; d < ndim; ++d) {
    const int64_t N = im_size[d], P = pad_size[d];
    if (N < 1 || P < N)
      return fail(TNL_EINVAL, "need 1 <= im_size <= pad_size on every axis");
    g.N[3 - ndim + d] = N;
    g.P[3 - ndim + d] = P;
  }
  if (nbatch == 0) return TNL_OK;
  if (!in || !out) return fail(TNL_EINVAL, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (weight) {
    switch (elem_bytes) {
      case 4: return launch_wrap<4, true>(pad, in, out, weight, g, s);
      case 8: return launch_wrap<8, true>(pad, in, out, weight, g, s);
      default:
        return fail(TNL_EINVAL, "a weight needs float32 or complex64 data (elem_bytes 4 or 8)");
    }
  }
  switch (elem_bytes) {
    case 4: return launch_wrap<4, false>(pad, in, out, nullptr, g, s);
    case 8: return launch_wrap<8, false>(pad, in, out, nullptr, g, s);
    case 16: return launch_wrap<16, false>(pad, in, out, nullptr, g, s);
    default: return fail(TNL_EINVAL, "elem_bytes must be 4, 8 or 16");
  }
}

}  // namespace

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_pad_wrapped(const void* x, void* y, int elem_bytes, int64_t nbatch,
                               int ndim, const int64_t* im_size,
                               const int64_t* pad_size, const float* weight,
                               void* stream) {
  return wrap_entry(true, x, y, elem_bytes, nbatch, ndim, im_size, pad_size, weight,
                    stream);
}

extern "C" int tnl_crop_wrapped(const void* y, void* x, int elem_bytes, int64_t nbatch,
                                int ndim, const int64_t* im_size,
                                const int64_t* pad_size, const float* weight,
                                void* stream) {
  return wrap_entry(false, y, x, elem_bytes, nbatch, ndim, im_size, pad_size, weight,
                    stream);
}
