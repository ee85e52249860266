// 3x3x3 convolutions ("same" zero padding) for the learned preconditioner.
//
// Activations are channel-last bf16, [nbatch, Z, Y, X, C] with C padded to
// a multiple of 8 (cin_p / cout_p).  Weights are one bf16 matrix
// W[16][Kp], row = output channel (rows >= cout are zero), column
// k = tap * cin_p + ci with tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1),
// zero-padded to Kp = 32 * ceil(27 cin_p / 32).
//
// Forward: implicit GEMM on the matrix cores.  A wave owns 64 consecutive
// voxels of one x row and all 16 output channels: per 32-wide K step it
// issues four v_mfma_f32_16x16x32_bf16 (one per 16-voxel group) that share
// the weight operand, which stays in registers for the whole kernel.  The
// input operand is gathered straight from memory (8 channels = 16 bytes
// per lane; the 27-fold reuse hits in cache).  Accumulation is fp32; bias
// and ReLU are fused into the store.  The data gradient is the same
// kernel run with flipped, channel-transposed weights.
//
// Weight gradient: dW[co][ci][tap] = sum over voxels of
// dout[co](p) * in[ci](p + tap); one workgroup per 8x8x8 voxel tile stages
// the tile (with halo) in LDS and writes a per-tile partial; a second
// kernel sums the partials in tile order (deterministic, no atomics).
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int CINP>
__global__ __launch_bounds__(256) void conv3d_fwd_kernel(
    const __bf16* __restrict__ in, const __bf16* __restrict__ w,
    const float* __restrict__ bias, __bf16* __restrict__ out_b,
    float* __restrict__ out_f, int64_t nrows, int Z, int Y, int X, int cout,
    int cout_p, int relu) {
  constexpr int KSTEPS = (27 * CINP + 31) / 32;
  const int lane = threadIdx.x & 63;
  const int xchunks = (X + 63) / 64;
  int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrows) return;  // whole wave; no barriers below
  const int xc = (int)(row % xchunks);
  row /= xchunks;
  const int yy = (int)(row % Y);
  row /= Y;
  const int zz = (int)(row % Z);
  const int64_t b = row / Z;
  const int x0 = xc * 64;
  const int kg = lane >> 4;  // K group: 8 consecutive k
  const int vl = lane & 15;  // voxel within a 16-group / output-channel row

  bf16x8 a[KSTEPS];
#pragma unroll
  for (int s = 0; s < KSTEPS; ++s)
    a[s] = *reinterpret_cast<const bf16x8*>(w + (int64_t)vl * (KSTEPS * 32) + s * 32 + kg * 8);

  f32x4 acc[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) acc[g] = f32x4{0.f, 0.f, 0.f, 0.f};
  bf16x8 zero;
#pragma unroll
  for (int j = 0; j < 8; ++j) zero[j] = (__bf16)0.0f;

#pragma unroll
  for (int s = 0; s < KSTEPS; ++s) {
    const int k0 = s * 32 + kg * 8;
    const int tap = k0 / CINP, ci0 = k0 - tap * CINP;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int sz = zz + dz, sy = yy + dy;
    const bool row_ok = tap < 27 && sz >= 0 && sz < Z && sy >= 0 && sy < Y;
    const int64_t rbase = row_ok ? ((b * Z + sz) * Y + sy) * (int64_t)X : 0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int sx = x0 + 16 * g + vl + dx;
      bf16x8 bv = zero;
      if (row_ok && sx >= 0 && sx < X)
        bv = *reinterpret_cast<const bf16x8*>(in + (rbase + sx) * CINP + ci0);
      acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s], bv, acc[g], 0, 0, 0);
    }
  }
  // lane holds output channels 4 kg + i of voxel x0 + 16 g + vl
  const int64_t obase = ((b * Z + zz) * Y + yy) * (int64_t)X;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int x = x0 + 16 * g + vl;
    if (x >= X) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = 4 * kg + i;
      if (co >= cout) continue;
      float v = acc[g][i] + (bias ? bias[co] : 0.0f);
      if (relu) v = fmaxf(v, 0.0f);
      const int64_t o = (obase + x) * cout_p + co;
      if (out_b) out_b[o] = (__bf16)v;
      if (out_f) out_f[o] = v;
    }
  }
}

constexpr int WG_T = 8;               // weight-gradient tile edge (voxels)
constexpr int WG_H = WG_T + 2;        // with halo

// Per-tile partial of dW for cout <= 16, cin_p <= 16: thread (co, ci)
// accumulates its 27 taps over the tile's 512 voxels.
__global__ __launch_bounds__(256) void conv3d_wgrad_partial_kernel(
    const __bf16* __restrict__ in, const __bf16* __restrict__ dout,
    float* __restrict__ partial, int Z, int Y, int X, int cin_p, int cout,
    int cout_p, int tz, int ty, int tx) {
  __shared__ __bf16 s_in[WG_H * WG_H * WG_H * 16];  // 32 KB
  __shared__ __bf16 s_d[WG_T * WG_T * WG_T * 16];   // 16 KB
  int t = blockIdx.x;
  const int bx = t % tx; t /= tx;
  const int by = t % ty; t /= ty;
  const int bz = t % tz;
  const int64_t b = t / tz;
  const int z0 = bz * WG_T, y0 = by * WG_T, x0 = bx * WG_T;
  for (int e = threadIdx.x; e < WG_H * WG_H * WG_H * cin_p; e += 256) {
    const int ci = e % cin_p, c = e / cin_p;
    const int hx = c % WG_H, hy = (c / WG_H) % WG_H, hz = c / (WG_H * WG_H);
    const int z = z0 + hz - 1, y = y0 + hy - 1, x = x0 + hx - 1;
    __bf16 v = (__bf16)0.0f;
    if (z >= 0 && z < Z && y >= 0 && y < Y && x >= 0 && x < X)
      v = in[(((b * Z + z) * Y + y) * (int64_t)X + x) * cin_p + ci];
    s_in[c * 16 + ci] = v;
  }
  for (int e = threadIdx.x; e < WG_T * WG_T * WG_T * 16; e += 256) {
    const int co = e % 16, c = e / 16;
    const int lx = c % WG_T, ly = (c / WG_T) % WG_T, lz = c / (WG_T * WG_T);
    const int z = z0 + lz, y = y0 + ly, x = x0 + lx;
    __bf16 v = (__bf16)0.0f;
    if (co < cout && z < Z && y < Y && x < X)
      v = dout[(((b * Z + z) * Y + y) * (int64_t)X + x) * cout_p + co];
    s_d[c * 16 + co] = v;
  }
  __syncthreads();
  const int co = threadIdx.x >> 4, ci = threadIdx.x & 15;
  float acc[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.0f;
  if (co < cout && ci < cin_p) {
    for (int c = 0; c < WG_T * WG_T * WG_T; ++c) {
      const float d = (float)s_d[c * 16 + co];
      const int lx = c % WG_T, ly = (c / WG_T) % WG_T, lz = c / (WG_T * WG_T);
#pragma unroll
      for (int k = 0; k < 27; ++k) {
        const int hz = lz + k / 9, hy = ly + (k / 3) % 3, hx = lx + k % 3;
        acc[k] = fmaf(d, (float)s_in[((hz * WG_H + hy) * WG_H + hx) * 16 + ci], acc[k]);
      }
    }
  }
  float* p = partial + (int64_t)blockIdx.x * 16 * 16 * 27 + (co * 16 + ci) * 27;
#pragma unroll
  for (int k = 0; k < 27; ++k) p[k] = acc[k];
}

__global__ void conv3d_wgrad_reduce_kernel(const float* __restrict__ partial,
                                           int64_t ntiles, float* __restrict__ dw,
                                           int cin_p, int cout) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // (co, ci, tap), 16 x 16 x 27
  if (e >= 16 * 16 * 27) return;
  const int co = e / (16 * 27), ci = (e / 27) % 16, k = e % 27;
  if (co >= cout || ci >= cin_p) return;
  float s = 0.0f;
  for (int64_t t = 0; t < ntiles; ++t) s += partial[t * 16 * 16 * 27 + e];
  dw[(co * cin_p + ci) * 27 + k] = s;
}

int conv3d_fwd(const void* in, const void* w, const float* bias, void* out_b,
               float* out_f, int64_t nbatch, int Z, int Y, int X, int cin_p,
               int cout, int cout_p, int relu, hipStream_t s) {
  if (nbatch < 1 || Z < 1 || Y < 1 || X < 1) return fail(TNL_EINVAL, "conv3d: empty volume");
  if (cout < 1 || cout > 16 || cout_p < cout || cout_p % 8)
    return fail(TNL_EINVAL, "conv3d: 1 <= cout <= 16, cout_p a multiple of 8 >= cout");
  if (!in || !w || (!out_b && !out_f)) return fail(TNL_EINVAL, "conv3d: null pointer");
  const int64_t nrows = nbatch * Z * Y * (int64_t)((X + 63) / 64);
  const unsigned blocks = (unsigned)((nrows + 3) / 4);
  const __bf16* ip = (const __bf16*)in;
  const __bf16* wp = (const __bf16*)w;
  __bf16* ob = (__bf16*)out_b;
  switch (cin_p) {
    case 8:
      conv3d_fwd_kernel<8><<<blocks, 256, 0, s>>>(ip, wp, bias, ob, out_f, nrows, Z, Y, X, cout, cout_p, relu);
      break;
    case 16:
      conv3d_fwd_kernel<16><<<blocks, 256, 0, s>>>(ip, wp, bias, ob, out_f, nrows, Z, Y, X, cout, cout_p, relu);
      break;
    default:
      return fail(TNL_EINVAL, "conv3d: cin_p must be 8 or 16");
  }
  TNL_LAUNCH_CHECK("conv3d_fwd");
  return TNL_OK;
}

int64_t conv3d_wgrad_ws_bytes(int64_t nbatch, int Z, int Y, int X) {
  const int64_t tiles = nbatch * (int64_t)((Z + WG_T - 1) / WG_T) * ((Y + WG_T - 1) / WG_T) *
                        ((X + WG_T - 1) / WG_T);
  return tiles * 16 * 16 * 27 * 4;
}

int conv3d_wgrad(const void* in, const void* dout, float* dw, float* ws,
                 int64_t ws_bytes, int64_t nbatch, int Z, int Y, int X, int cin_p,
                 int cout, int cout_p, hipStream_t s) {
  if (nbatch < 1 || Z < 1 || Y < 1 || X < 1) return fail(TNL_EINVAL, "conv3d: empty volume");
  if (cin_p != 8 && cin_p != 16) return fail(TNL_EINVAL, "conv3d: cin_p must be 8 or 16");
  if (cout < 1 || cout > 16 || cout_p < cout || cout_p % 8)
    return fail(TNL_EINVAL, "conv3d: 1 <= cout <= 16, cout_p a multiple of 8 >= cout");
  if (!in || !dout || !dw || !ws) return fail(TNL_EINVAL, "conv3d: null pointer");
  if (ws_bytes < conv3d_wgrad_ws_bytes(nbatch, Z, Y, X))
    return fail(TNL_EWORKSPACE, "conv3d_wgrad: workspace too small");
  const int tz = (Z + WG_T - 1) / WG_T, ty = (Y + WG_T - 1) / WG_T, tx = (X + WG_T - 1) / WG_T;
  const int64_t ntiles = nbatch * tz * ty * (int64_t)tx;
  conv3d_wgrad_partial_kernel<<<(unsigned)ntiles, 256, 0, s>>>(
      (const __bf16*)in, (const __bf16*)dout, ws, Z, Y, X, cin_p, cout, cout_p, tz, ty, tx);
  TNL_LAUNCH_CHECK("conv3d_wgrad_partial");
  conv3d_wgrad_reduce_kernel<<<(16 * 16 * 27 + 255) / 256, 256, 0, s>>>(ws, ntiles, dw, cin_p,
                                                                       cout);
  TNL_LAUNCH_CHECK("conv3d_wgrad_reduce");
  return TNL_OK;
}

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_conv3d_fwd(const void* in, const void* w, const float* bias,
                              void* out_bf16, float* out_f32, int64_t nbatch, int z,
                              int y, int x, int cin_p, int cout, int cout_p, int relu,
                              void* stream) {
  return conv3d_fwd(in, w, bias, out_bf16, out_f32, nbatch, z, y, x, cin_p, cout, cout_p,
                    relu, (hipStream_t)stream);
}

extern "C" int64_t tnl_conv3d_wgrad_workspace(int64_t nbatch, int z, int y, int x) {
  return conv3d_wgrad_ws_bytes(nbatch, z, y, x);
}

extern "C" int tnl_conv3d_wgrad(const void* in, const void* dout, float* dw, float* ws,
                                int64_t ws_bytes, int64_t nbatch, int z, int y, int x,
                                int cin_p, int cout, int cout_p, void* stream) {
  return conv3d_wgrad(in, dout, dw, ws, ws_bytes, nbatch, z, y, x, cin_p, cout, cout_p,
                      (hipStream_t)stream);
}
