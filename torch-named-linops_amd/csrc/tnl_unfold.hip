// array_to_blocks (unfold) and blocks_to_array (fold) for gfx950.
//
// Both are HBM-bound data movement.  Dimensions are left-padded to 3 with
// size-1 axes.  The last axis is moved in vectors of V elements (16 bytes
// when the sizes, strides and pointers allow it) so every lane issues
// dwordx4 loads/stores.
//
// unfold: output-major copy, 32-bit decode of each output vector when the
//         sizes allow (UN vectors in flight per thread), else block-major:
//         a workgroup walks whole blocks and each lane owns fixed element
//         slots inside a block, its in-block offsets computed once.
// fold:   output-major gather.  Each lane owns V consecutive output elements
//         and adds the covering blocks' values in lexicographic block order,
//         which reproduces the reference's sequential `out[slc] += x[blk]`
//         loop (src/torchlinops/functional/_unfold/fold.py:508-516) bit for
//         bit, with no atomics.
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

struct BlockGeom {
  int64_t nb[3], B[3], im[3], s[3];  // last dim in vectors
  int64_t nbatch;
};

template <int VB>
struct VecT;
template <>
struct VecT<16> { using T = uint4; };
template <>
struct VecT<8> { using T = uint2; };
template <>
struct VecT<4> { using T = uint32_t; };
template <>
struct VecT<2> { using T = uint16_t; };
template <>
struct VecT<1> { using T = uint8_t; };

constexpr int UNFOLD_THREADS = 256;
constexpr int UNFOLD_SLOTS = 8;  // element slots per lane kept in registers

template <typename T>
__global__ __launch_bounds__(UNFOLD_THREADS) void unfold_kernel(
    const T* __restrict__ x, T* __restrict__ y, BlockGeom g) {
  const int64_t B12 = g.B[1] * g.B[2];
  const int64_t vpb = g.B[0] * B12;  // vectors per block
  const int64_t S1 = g.im[2];
  const int64_t S0 = g.im[1] * S1;
  const int64_t img = g.im[0] * S0;
  int64_t ioff[UNFOLD_SLOTS];
#pragma unroll
  for (int k = 0; k < UNFOLD_SLOTS; ++k) {
    const int64_t e = threadIdx.x + (int64_t)k * UNFOLD_THREADS;
    const int64_t j0 = e / B12;
    const int64_t r = e - j0 * B12;
    const int64_t j1 = r / g.B[2];
    const int64_t j2 = r - j1 * g.B[2];
    ioff[k] = j0 * S0 + j1 * S1 + j2;
  }
  const int64_t nb12 = g.nb[1] * g.nb[2];
  const int64_t nbb = g.nb[0] * nb12;
  const int64_t total = g.nbatch * nbb;
  for (int64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    const int64_t n = blk / nbb;
    int64_t r = blk - n * nbb;
    const int64_t b0 = r / nb12;
    r -= b0 * nb12;
    const int64_t b1 = r / g.nb[2];
    const int64_t b2 = r - b1 * g.nb[2];
    const int64_t ibase =
        n * img + b0 * g.s[0] * S0 + b1 * g.s[1] * S1 + b2 * g.s[2];
    const int64_t obase = blk * vpb;
#pragma unroll
    for (int k = 0; k < UNFOLD_SLOTS; ++k) {
      const int64_t e = threadIdx.x + (int64_t)k * UNFOLD_THREADS;
      if (e < vpb) stream_store(&y[obase + e], x[ibase + ioff[k]]);
    }
  }
}

// Blocks too large for the register slots: offsets computed per element.
template <typename T>
__global__ __launch_bounds__(UNFOLD_THREADS) void unfold_large_kernel(
    const T* __restrict__ x, T* __restrict__ y, BlockGeom g) {
  const int64_t B12 = g.B[1] * g.B[2];
  const int64_t vpb = g.B[0] * B12;
  const int64_t S1 = g.im[2];
  const int64_t S0 = g.im[1] * S1;
  const int64_t img = g.im[0] * S0;
  const int64_t nb12 = g.nb[1] * g.nb[2];
  const int64_t nbb = g.nb[0] * nb12;
  const int64_t total = g.nbatch * nbb;
  for (int64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    const int64_t n = blk / nbb;
    int64_t r = blk - n * nbb;
    const int64_t b0 = r / nb12;
    r -= b0 * nb12;
    const int64_t b1 = r / g.nb[2];
    const int64_t b2 = r - b1 * g.nb[2];
    const int64_t ibase =
        n * img + b0 * g.s[0] * S0 + b1 * g.s[1] * S1 + b2 * g.s[2];
    const int64_t obase = blk * vpb;
    for (int64_t e = threadIdx.x; e < vpb; e += UNFOLD_THREADS) {
      const int64_t j0 = e / B12;
      const int64_t q = e - j0 * B12;
      const int64_t j1 = q / g.B[2];
      const int64_t j2 = q - j1 * g.B[2];
      stream_store(&y[obase + e], x[ibase + j0 * S0 + j1 * S1 + j2]);
    }
  }
}

// Covering block range of output index i along one axis:
// blocks b with b*s <= i < b*s + B and 0 <= b < nb.
__device__ __forceinline__ void cover(int64_t i, int64_t B, int64_t s,
                                      int64_t nb, int64_t& lo, int64_t& hi) {
  const int64_t a = i - B + 1;
  lo = a <= 0 ? 0 : (a + s - 1) / s;
  hi = i / s;
  if (hi > nb - 1) hi = nb - 1;
}

template <typename S, int V>
struct Pack {
  S v[V];
};

template <typename S, int V>
__global__ __launch_bounds__(256) void fold_kernel(const S* __restrict__ x,
                                                   S* __restrict__ y,
                                                   BlockGeom g) {
  using P = Pack<S, V>;
  const P* __restrict__ xv = reinterpret_cast<const P*>(x);
  P* __restrict__ yv = reinterpret_cast<P*>(y);
  const int64_t S1 = g.im[2];
  const int64_t S0 = g.im[1] * S1;
  const int64_t img = g.im[0] * S0;
  const int64_t total = g.nbatch * img;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += stride) {
    const int64_t n = q / img;
    int64_t r = q - n * img;
    const int64_t i0 = r / S0;
    r -= i0 * S0;
    const int64_t i1 = r / S1;
    const int64_t i2 = r - i1 * S1;
    int64_t lo0, hi0, lo1, hi1, lo2, hi2;
    cover(i0, g.B[0], g.s[0], g.nb[0], lo0, hi0);
    cover(i1, g.B[1], g.s[1], g.nb[1], lo1, hi1);
    cover(i2, g.B[2], g.s[2], g.nb[2], lo2, hi2);
    P acc;
#pragma unroll
    for (int c = 0; c < V; ++c) acc.v[c] = (S)0;
    const int64_t bsz = g.B[0] * g.B[1] * g.B[2];
    for (int64_t b0 = lo0; b0 <= hi0; ++b0) {
      for (int64_t b1 = lo1; b1 <= hi1; ++b1) {
        const int64_t row =
            ((n * g.nb[0] + b0) * g.nb[1] + b1) * g.nb[2];
        const int64_t inblk01 =
            ((i0 - b0 * g.s[0]) * g.B[1] + (i1 - b1 * g.s[1])) * g.B[2];
        for (int64_t b2 = lo2; b2 <= hi2; ++b2) {
          const P v = xv[(row + b2) * bsz + inblk01 + (i2 - b2 * g.s[2])];
#pragma unroll
          for (int c = 0; c < V; ++c) acc.v[c] += v.v[c];
        }
      }
    }
    yv[q] = acc;
  }
}

// ---------------------------------------------------------------------------
// 32-bit variants, taken when every element index of the input and output
// is below 2^31 (one 8 x 128^3 complex64 image and its 31^3 blocks of 8^3
// are 1.3e8 elements).  Indices are decoded with divisions by invariant
// divisors done as a multiply-high and a shift (magic numbers computed on
// the host, exact for every numerator below 2^31), where the 64-bit kernels
// above divide in software.  Each thread moves UN vectors per iteration,
// spaced a grid apart, so UN loads are in flight before the stores.

struct FDiv {
  uint32_t d, m, l;  // n / d == (mulhi(n, m) + n) >> l for n < 2^31
};

static FDiv make_fdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
  return FDiv{d, (uint32_t)m, l};
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FDiv& f) {
  return (uint32_t)(((uint64_t)__umulhi(n, f.m) + n) >> f.l);
}

struct BlockGeom32 {
  FDiv vpb, B12, B2, nbb, nb12, nb2;  // unfold decode
  FDiv img, S0, S1, s[3];             // fold decode
  uint32_t B[3], nb[3], im[3], st[3];
  uint32_t S0v, S1v, imgv, bsz;
  uint32_t total;  // vectors of the output
};

constexpr int UN = 4;  // vectors per thread per iteration

template <typename T>
__device__ __forceinline__ T unfold32_load(const T* __restrict__ x, const BlockGeom32& g,
                                           uint32_t o) {
  const uint32_t blk = fdiv(o, g.vpb);
  const uint32_t e = o - blk * g.vpb.d;
  const uint32_t j0 = fdiv(e, g.B12);
  const uint32_t r = e - j0 * g.B12.d;
  const uint32_t j1 = fdiv(r, g.B2);
  const uint32_t j2 = r - j1 * g.B2.d;
  const uint32_t n = fdiv(blk, g.nbb);
  uint32_t rb = blk - n * g.nbb.d;
  const uint32_t b0 = fdiv(rb, g.nb12);
  rb -= b0 * g.nb12.d;
  const uint32_t b1 = fdiv(rb, g.nb2);
  const uint32_t b2 = rb - b1 * g.nb2.d;
  return x[n * g.imgv + (b0 * g.st[0] + j0) * g.S0v + (b1 * g.st[1] + j1) * g.S1v +
           b2 * g.st[2] + j2];
}

// Each workgroup copies one contiguous stretch of the output (chunk vectors,
// UN per thread in flight), and workgroup ids are remapped so the hardware's
// round-robin over the 8 XCDs hands each XCD a contiguous run of stretches:
// blocks that overlap in the first two axes (31 blocks and 31^2 blocks
// apart in the output at 31 blocks per axis) then run on the same XCD close
// in time and find their shared image rows in its L2.
template <typename T>
__global__ __launch_bounds__(256) void unfold32_kernel(const T* __restrict__ x,
                                                       T* __restrict__ y, BlockGeom32 g,
                                                       uint32_t chunk) {
  const uint32_t nwg = gridDim.x;  // a multiple of 8
  const uint32_t w = (blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8;
  const uint64_t lo64 = (uint64_t)w * chunk;
  if (lo64 >= g.total) return;
  const uint32_t lo = (uint32_t)lo64;
  const uint32_t hi = g.total - lo < chunk ? g.total : lo + chunk;
  for (uint32_t o0 = lo + threadIdx.x; o0 < hi; o0 += 256u * UN) {
    T v[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t o = o0 + 256u * u;
      if (o < hi) v[u] = unfold32_load(x, g, o);
    }
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      const uint32_t o = o0 + 256u * u;
      if (o < hi) stream_store(&y[o], v[u]);
    }
  }
}

// Covering blocks of output index i along one axis (32-bit).
__device__ __forceinline__ void cover32(uint32_t i, uint32_t B, const FDiv& s, uint32_t nb,
                                        uint32_t& lo, uint32_t& hi) {
  lo = i + 1 <= B ? 0u : fdiv(i + 1 - B + s.d - 1, s);
  hi = fdiv(i, s);
  if (hi > nb - 1) hi = nb - 1;
}

template <typename S, int V>
__global__ __launch_bounds__(256) void fold32_kernel(const S* __restrict__ x,
                                                     S* __restrict__ y, BlockGeom32 g) {
  using P = Pack<S, V>;
  const P* __restrict__ xv = reinterpret_cast<const P*>(x);
  P* __restrict__ yv = reinterpret_cast<P*>(y);
  const uint32_t stride = gridDim.x * 256u;
  for (uint32_t q = blockIdx.x * 256u + threadIdx.x; q < g.total; q += stride) {
    const uint32_t n = fdiv(q, g.img);
    uint32_t r = q - n * g.imgv;
    const uint32_t i0 = fdiv(r, g.S0);
    r -= i0 * g.S0v;
    const uint32_t i1 = fdiv(r, g.S1);
    const uint32_t i2 = r - i1 * g.S1v;
    uint32_t lo0, hi0, lo1, hi1, lo2, hi2;
    cover32(i0, g.B[0], g.s[0], g.nb[0], lo0, hi0);
    cover32(i1, g.B[1], g.s[1], g.nb[1], lo1, hi1);
    cover32(i2, g.B[2], g.s[2], g.nb[2], lo2, hi2);
    P acc;
#pragma unroll
    for (int c = 0; c < V; ++c) acc.v[c] = (S)0;
    for (uint32_t b0 = lo0; b0 <= hi0; ++b0) {
      for (uint32_t b1 = lo1; b1 <= hi1; ++b1) {
        const uint32_t row = ((n * g.nb[0] + b0) * g.nb[1] + b1) * g.nb[2];
        const uint32_t inblk01 =
            ((i0 - b0 * g.st[0]) * g.B[1] + (i1 - b1 * g.st[1])) * g.B[2];
        for (uint32_t b2 = lo2; b2 <= hi2; ++b2) {
          const P v = xv[(row + b2) * g.bsz + inblk01 + (i2 - b2 * g.st[2])];
#pragma unroll
          for (int c = 0; c < V; ++c) acc.v[c] += v.v[c];
        }
      }
    }
    stream_store(&yv[q], acc);
  }
}

// The 32-bit geometry when every index fits, else false.
static bool make_geom32(const BlockGeom& g, bool fold, BlockGeom32& h) {
  const int64_t lim = (int64_t)1 << 31;
  const int64_t nblk = g.nbatch * g.nb[0] * g.nb[1] * g.nb[2];
  const int64_t vpb = g.B[0] * g.B[1] * g.B[2];
  const int64_t out_unfold = nblk * vpb;
  const int64_t img = g.im[0] * g.im[1] * g.im[2];
  if (out_unfold >= lim || g.nbatch * img >= lim) return false;
  h.vpb = make_fdiv((uint32_t)vpb);
  h.B12 = make_fdiv((uint32_t)(g.B[1] * g.B[2]));
  h.B2 = make_fdiv((uint32_t)g.B[2]);
  h.nbb = make_fdiv((uint32_t)(g.nb[0] * g.nb[1] * g.nb[2]));
  h.nb12 = make_fdiv((uint32_t)(g.nb[1] * g.nb[2]));
  h.nb2 = make_fdiv((uint32_t)g.nb[2]);
  h.img = make_fdiv((uint32_t)img);
  h.S0 = make_fdiv((uint32_t)(g.im[1] * g.im[2]));
  h.S1 = make_fdiv((uint32_t)g.im[2]);
  for (int d = 0; d < 3; ++d) {
    h.s[d] = make_fdiv((uint32_t)g.s[d]);
    h.B[d] = (uint32_t)g.B[d];
    h.nb[d] = (uint32_t)g.nb[d];
    h.im[d] = (uint32_t)g.im[d];
    h.st[d] = (uint32_t)g.s[d];
  }
  h.S0v = (uint32_t)(g.im[1] * g.im[2]);
  h.S1v = (uint32_t)g.im[2];
  h.imgv = (uint32_t)img;
  h.bsz = (uint32_t)vpb;
  h.total = (uint32_t)(fold ? g.nbatch * img : out_unfold);
  return true;
}

static unsigned grid_for(uint32_t total, int per_thread) {
  int64_t wg = ((int64_t)total + 256 * per_thread - 1) / (256 * per_thread);
  if (wg > 65536) wg = 65536;
  return (unsigned)(wg < 1 ? 1 : wg);
}

static void pad3(int ndim, const int64_t* a, int64_t fill, int64_t out[3]) {
  for (int d = 0; d < 3; ++d) out[d] = fill;
  for (int d = 0; d < ndim; ++d) out[3 - ndim + d] = a[d];
}

static int check_block_args(int ndim, const int64_t* nblocks,
                            const int64_t* block_size, const int64_t* im_size,
                            const int64_t* stride) {
  if (ndim < 1 || ndim > 3) return fail(TNL_EINVAL, "ndim must be 1, 2 or 3");
  for (int d = 0; d < ndim; ++d) {
    if (block_size[d] < 1 || stride[d] < 1 || im_size[d] < 1)
      return fail(TNL_EINVAL, "block_size, stride and im_size must be >= 1");
    if (nblocks[d] < 1)
      return fail(TNL_EINVAL, "Found 0 in nblocks - make sure there is at "
                              "least one block in each direction.");
    if ((im_size[d] - block_size[d]) / stride[d] + 1 != nblocks[d])
      return fail(TNL_ESHAPE, "nblocks inconsistent with im_size, "
                              "block_size and stride");
  }
  return TNL_OK;
}

// Largest vector width (bytes) usable for the last axis.
static int pick_vec_bytes(int elem_bytes, const int64_t B[3],
                          const int64_t im[3], const int64_t s[3],
                          const void* x, const void* y) {
  for (int vb = 16; vb >= elem_bytes; vb /= 2) {
    const int64_t v = vb / elem_bytes;
    if (B[2] % v || im[2] % v || s[2] % v) continue;
    if (((uintptr_t)x | (uintptr_t)y) % vb) continue;
    return vb;
  }
  return elem_bytes;
}

template <int VB>
static int launch_unfold(const void* x, void* y, BlockGeom g, hipStream_t st) {
  using T = typename VecT<VB>::T;
  BlockGeom32 h;
  if (make_geom32(g, false, h)) {
    // stretches of 256 x UN vectors (one pass of the workgroup; measured on
    // 8 x 128^3 complex64, 8^3 blocks, stride 4: 284 us against 295 us for
    // twice that and 311 us for a quarter), a multiple of 8 workgroups
    const uint32_t chunk = 256u * UN;
    int64_t nwg = ((int64_t)h.total + chunk - 1) / chunk;
    nwg = (nwg + 7) / 8 * 8;
    unfold32_kernel<T><<<(unsigned)nwg, 256, 0, st>>>((const T*)x, (T*)y, h, chunk);
    TNL_LAUNCH_CHECK("unfold");
    return TNL_OK;
  }
  const int64_t vpb = g.B[0] * g.B[1] * g.B[2];
  const int64_t total = g.nbatch * g.nb[0] * g.nb[1] * g.nb[2];
  int64_t grid = total < 16384 ? total : 16384;
  if (grid < 1) grid = 1;
  if (vpb <= (int64_t)UNFOLD_SLOTS * UNFOLD_THREADS)
    unfold_kernel<T><<<(unsigned)grid, UNFOLD_THREADS, 0, st>>>(
        (const T*)x, (T*)y, g);
  else
    unfold_large_kernel<T><<<(unsigned)grid, UNFOLD_THREADS, 0, st>>>(
        (const T*)x, (T*)y, g);
  TNL_LAUNCH_CHECK("unfold");
  return TNL_OK;
}

template <typename S, int V>
static int launch_fold(const void* x, void* y, BlockGeom g, hipStream_t st) {
  BlockGeom32 h;
  if (make_geom32(g, true, h)) {
    fold32_kernel<S, V><<<grid_for(h.total, 1), 256, 0, st>>>((const S*)x, (S*)y, h);
    TNL_LAUNCH_CHECK("fold");
    return TNL_OK;
  }
  const int64_t total = g.nbatch * g.im[0] * g.im[1] * g.im[2];
  int64_t grid = (total + 255) / 256;
  if (grid > 32768) grid = 32768;
  if (grid < 1) grid = 1;
  fold_kernel<S, V><<<(unsigned)grid, 256, 0, st>>>((const S*)x, (S*)y, g);
  TNL_LAUNCH_CHECK("fold");
  return TNL_OK;
}

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_unfold(const void* x, void* y, int elem_bytes,
                          int64_t nbatch, int ndim, const int64_t* nblocks,
                          const int64_t* block_size, const int64_t* im_size,
                          const int64_t* stride, void* stream) {
  if (int rc = check_block_args(ndim, nblocks, block_size, im_size, stride))
    return rc;
  if (!(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 ||
        elem_bytes == 8))
    return fail(TNL_EINVAL, "elem_bytes must be 1, 2, 4 or 8");
  if (nbatch < 0) return fail(TNL_EINVAL, "nbatch must be >= 0");
  if (nbatch == 0) return TNL_OK;
  if (!x || !y) return fail(TNL_EINVAL, "null pointer");
  BlockGeom g;
  pad3(ndim, nblocks, 1, g.nb);
  pad3(ndim, block_size, 1, g.B);
  pad3(ndim, im_size, 1, g.im);
  pad3(ndim, stride, 1, g.s);
  g.nbatch = nbatch;
  const int vb = pick_vec_bytes(elem_bytes, g.B, g.im, g.s, x, y);
  const int64_t v = vb / elem_bytes;
  g.B[2] /= v;
  g.im[2] /= v;
  g.s[2] /= v;
  hipStream_t st = (hipStream_t)stream;
  switch (vb) {
    case 16: return launch_unfold<16>(x, y, g, st);
    case 8: return launch_unfold<8>(x, y, g, st);
    case 4: return launch_unfold<4>(x, y, g, st);
    case 2: return launch_unfold<2>(x, y, g, st);
    default: return launch_unfold<1>(x, y, g, st);
  }
}

extern "C" int tnl_fold(const void* x, void* y, int dtype, int64_t nbatch,
                        int ndim, const int64_t* nblocks,
                        const int64_t* block_size, const int64_t* im_size,
                        const int64_t* stride, void* stream) {
  if (int rc = check_block_args(ndim, nblocks, block_size, im_size, stride))
    return rc;
  if (dtype != TNL_F32 && dtype != TNL_F64)
    return fail(TNL_EINVAL, "fold dtype must be float32 or float64");
  if (nbatch < 0) return fail(TNL_EINVAL, "nbatch must be >= 0");
  if (nbatch == 0) return TNL_OK;
  if (!x || !y) return fail(TNL_EINVAL, "null pointer");
  BlockGeom g;
  pad3(ndim, nblocks, 1, g.nb);
  pad3(ndim, block_size, 1, g.B);
  pad3(ndim, im_size, 1, g.im);
  pad3(ndim, stride, 1, g.s);
  g.nbatch = nbatch;
  const int eb = dtype == TNL_F32 ? 4 : 8;
  const int vb = pick_vec_bytes(eb, g.B, g.im, g.s, x, y);
  const int64_t v = vb / eb;
  g.B[2] /= v;
  g.im[2] /= v;
  g.s[2] /= v;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TNL_F32) {
    switch (v) {
      case 4: return launch_fold<float, 4>(x, y, g, st);
      case 2: return launch_fold<float, 2>(x, y, g, st);
      default: return launch_fold<float, 1>(x, y, g, st);
    }
  }
  switch (v) {
    case 2: return launch_fold<double, 2>(x, y, g, st);
    default: return launch_fold<double, 1>(x, y, g, st);
  }
}
