// Segmented vector kernels for solvers that run many independent problems
// over one flat float buffer (complex data as interleaved floats: the real
// part of conj(a) . b is the plain dot product of the float views).
//
// Segment s covers floats [start_s, start_s + len_s).  The buffer is cut
// into reduction blocks that never straddle a segment; a block writes one
// float64 partial, and a second kernel sums each segment's partials in
// block order.  No atomics and no memsets: results are deterministic and
// the kernels can be captured in a graph and replayed.
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

#include <vector>

namespace tnl {

constexpr int SEG_THREADS = 256;

struct SegBlock {
  int64_t begin, end;  // float range of the block
};

// Block-wide float64 sum (256 threads), result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[SEG_THREADS / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < SEG_THREADS / 64; ++w) t += red[w];
  return t;
}

__global__ __launch_bounds__(SEG_THREADS) void seg_dot_kernel(
    const float* __restrict__ a, const float* __restrict__ b,
    const SegBlock* __restrict__ blocks, double* __restrict__ partial) {
  const SegBlock bl = blocks[blockIdx.x];
  double acc = 0.0;
  for (int64_t i = bl.begin + threadIdx.x; i < bl.end; i += SEG_THREADS)
    acc += (double)a[i] * (double)b[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void seg_reduce_kernel(const double* __restrict__ partial,
                                  const int32_t* __restrict__ segblk, int nseg,
                                  double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  double t = 0.0;
  for (int b = segblk[s]; b < segblk[s + 1]; ++b) t += partial[b];
  out[s] = t;
}

// x += alpha p, r -= alpha Ap with alpha = rs / pAp of the block's segment
// (0 when pAp is 0), and the block's partial of |r|^2 after the update.
__global__ __launch_bounds__(SEG_THREADS) void seg_cg_update_kernel(
    float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p,
    const float* __restrict__ Ap, const double* __restrict__ rs,
    const double* __restrict__ pAp, const SegBlock* __restrict__ blocks,
    const int32_t* __restrict__ blkseg, double* __restrict__ partial) {
  const SegBlock bl = blocks[blockIdx.x];
  const int s = blkseg[blockIdx.x];
  const double den = pAp[s];
  const float alpha = den != 0.0 ? (float)(rs[s] / den) : 0.0f;
  double acc = 0.0;
  for (int64_t i = bl.begin + threadIdx.x; i < bl.end; i += SEG_THREADS) {
    x[i] = fmaf(alpha, p[i], x[i]);
    const float ri = fmaf(-alpha, Ap[i], r[i]);
    r[i] = ri;
    acc += (double)ri * (double)ri;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = r + beta p with beta = rs_new / rs of the block's segment (0 when
// rs is 0).
__global__ __launch_bounds__(SEG_THREADS) void seg_cg_direction_kernel(
    float* __restrict__ p, const float* __restrict__ r,
    const double* __restrict__ rs, const double* __restrict__ rs_new,
    const SegBlock* __restrict__ blocks, const int32_t* __restrict__ blkseg) {
  const SegBlock bl = blocks[blockIdx.x];
  const int s = blkseg[blockIdx.x];
  const double den = rs[s];
  const float beta = den != 0.0 ? (float)(rs_new[s] / den) : 0.0f;
  for (int64_t i = bl.begin + threadIdx.x; i < bl.end; i += SEG_THREADS)
    p[i] = fmaf(beta, p[i], r[i]);
}

__global__ void seg_copy_kernel(double* __restrict__ dst,
                                const double* __restrict__ src, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int seg_layout(const int64_t* seg_len, int nseg, int64_t block_floats,
               tnl_seg_layout* L) {
  if (nseg < 1 || !seg_len) return fail(TNL_EINVAL, "need at least one segment");
  if (block_floats <= 0) block_floats = 16384;
  std::memset(L, 0, sizeof(*L));
  L->nseg = nseg;
  L->block_floats = block_floats;
  int64_t nblocks = 0, numel = 0;
  for (int s = 0; s < nseg; ++s) {
    if (seg_len[s] < 0) return fail(TNL_EINVAL, "segment lengths must be >= 0");
    nblocks += seg_len[s] == 0 ? 0 : (seg_len[s] + block_floats - 1) / block_floats;
    numel += seg_len[s];
  }
  if (nblocks > (int64_t)1 << 30) return fail(TNL_EINVAL, "too many reduction blocks");
  L->numel = numel;
  L->nblocks = nblocks;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  };
  L->off_blocks = take(16 * std::max<int64_t>(nblocks, 1));
  L->off_blkseg = take(4 * std::max<int64_t>(nblocks, 1));
  L->off_segblk = take(4 * (nseg + 1));
  L->off_partial = take(8 * std::max<int64_t>(nblocks, 1));
  L->bytes = off;
  return TNL_OK;
}

int seg_build(const int64_t* seg_len, const tnl_seg_layout& L, void* ws,
              hipStream_t s) {
  std::vector<char> host((size_t)L.off_partial, 0);
  SegBlock* blocks = (SegBlock*)(host.data() + L.off_blocks);
  int32_t* blkseg = (int32_t*)(host.data() + L.off_blkseg);
  int32_t* segblk = (int32_t*)(host.data() + L.off_segblk);
  int64_t b = 0, start = 0;
  for (int sg = 0; sg < L.nseg; ++sg) {
    segblk[sg] = (int32_t)b;
    for (int64_t o = 0; o < seg_len[sg]; o += L.block_floats, ++b) {
      blocks[b].begin = start + o;
      blocks[b].end = start + std::min(seg_len[sg], o + L.block_floats);
      blkseg[b] = sg;
    }
    start += seg_len[sg];
  }
  segblk[L.nseg] = (int32_t)b;
  int rc = check_hip(hipMemcpyAsync(ws, host.data(), host.size(), hipMemcpyHostToDevice, s),
                     "segment table copy");
  if (!rc) rc = check_hip(hipStreamSynchronize(s), "segment table copy");
  return rc;
}

static int seg_reduce(const tnl_seg_layout& L, const char* ws, double* out, hipStream_t s) {
  seg_reduce_kernel<<<(L.nseg + 255) / 256, 256, 0, s>>>(
      (const double*)(ws + L.off_partial), (const int32_t*)(ws + L.off_segblk), L.nseg, out);
  TNL_LAUNCH_CHECK("seg_reduce");
  return TNL_OK;
}

int seg_dot(const float* a, const float* b, double* out, const tnl_seg_layout& L,
            void* ws, hipStream_t s) {
  char* w = (char*)ws;
  if (L.nblocks > 0) {
    seg_dot_kernel<<<(unsigned)L.nblocks, SEG_THREADS, 0, s>>>(
        a, b, (const SegBlock*)(w + L.off_blocks), (double*)(w + L.off_partial));
    TNL_LAUNCH_CHECK("seg_dot");
  }
  return seg_reduce(L, w, out, s);
}

int seg_cg_update(float* x, float* r, const float* p, const float* Ap,
                  const double* rs, const double* pAp, double* rs_new,
                  const tnl_seg_layout& L, void* ws, hipStream_t s) {
  char* w = (char*)ws;
  if (L.nblocks > 0) {
    seg_cg_update_kernel<<<(unsigned)L.nblocks, SEG_THREADS, 0, s>>>(
        x, r, p, Ap, rs, pAp, (const SegBlock*)(w + L.off_blocks),
        (const int32_t*)(w + L.off_blkseg), (double*)(w + L.off_partial));
    TNL_LAUNCH_CHECK("seg_cg_update");
  }
  return seg_reduce(L, w, rs_new, s);
}

int seg_cg_direction(float* p, const float* r, double* rs, const double* rs_new,
                     const tnl_seg_layout& L, void* ws, hipStream_t s) {
  char* w = (char*)ws;
  if (L.nblocks > 0) {
    seg_cg_direction_kernel<<<(unsigned)L.nblocks, SEG_THREADS, 0, s>>>(
        p, r, rs, rs_new, (const SegBlock*)(w + L.off_blocks),
        (const int32_t*)(w + L.off_blkseg));
    TNL_LAUNCH_CHECK("seg_cg_direction");
  }
  seg_copy_kernel<<<(L.nseg + 255) / 256, 256, 0, s>>>(rs, rs_new, L.nseg);
  TNL_LAUNCH_CHECK("seg_copy");
  return TNL_OK;
}

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_seg_layout_query(const int64_t* seg_len, int nseg,
                                    int64_t block_floats, tnl_seg_layout* layout) {
  if (!layout) return fail(TNL_EINVAL, "null layout");
  return seg_layout(seg_len, nseg, block_floats, layout);
}

extern "C" int tnl_seg_build(const int64_t* seg_len, const tnl_seg_layout* layout,
                             void* ws, void* stream) {
  if (!seg_len || !layout || !ws) return fail(TNL_EINVAL, "null pointer");
  return seg_build(seg_len, *layout, ws, (hipStream_t)stream);
}

extern "C" int tnl_seg_dot(const float* a, const float* b, double* out,
                           const tnl_seg_layout* layout, void* ws, void* stream) {
  if (!a || !b || !out || !layout || !ws) return fail(TNL_EINVAL, "null pointer");
  return seg_dot(a, b, out, *layout, ws, (hipStream_t)stream);
}

extern "C" int tnl_seg_cg_update(float* x, float* r, const float* p, const float* Ap,
                                 const double* rs, const double* pAp, double* rs_new,
                                 const tnl_seg_layout* layout, void* ws, void* stream) {
  if (!x || !r || !p || !Ap || !rs || !pAp || !rs_new || !layout || !ws)
    return fail(TNL_EINVAL, "null pointer");
  return seg_cg_update(x, r, p, Ap, rs, pAp, rs_new, *layout, ws, (hipStream_t)stream);
}

extern "C" int tnl_seg_cg_direction(float* p, const float* r, double* rs,
                                    const double* rs_new, const tnl_seg_layout* layout,
                                    void* ws, void* stream) {
  if (!p || !r || !rs || !rs_new || !layout || !ws) return fail(TNL_EINVAL, "null pointer");
  return seg_cg_direction(p, r, rs, rs_new, *layout, ws, (hipStream_t)stream);
}
