// Integer-index sampling for gfx950: the gather behind index / Sampling.fn
// and its adjoint, the scatter-add behind index_adjoint / Sampling.adj_fn.
//
// Both are HBM-bound byte movement over a flattened grid [nbatch, ngrid]
// and a point list [nbatch, npts].
//
// gather:  point-major.  A lane owns one point, loads its cell index once
//          and copies that cell for IB consecutive batch entries, so the
//          output rows are written coalesced and the index is read once
//          per IB batch entries.
// adjoint: cell-major gather over a CSR map (points grouped by cell in
//          ascending point order: `order`, `starts`).  A lane owns one
//          cell, sums its points in ascending order from zero (loads
//          issued PF at a time, adds kept in order) and writes the cell
//          for IB batch entries; cells that no point hits are written as
//          zero.  Every cell is written once with no atomics, so the
//          result does not depend on scheduling and equals the sequential
//          `y[idx[p]] += x[p]` loop of index_add_ on the CPU bit for bit.
//          Cells holding more than `heavy_min` points (the k-space centre
//          of a rounded radial trajectory holds thousands) are skipped
//          there and summed by a workgroup each: 256 contiguous sequential
//          partial sums combined by a fixed LDS tree, so they stay
//          deterministic but round differently from the sequential sum.
//          Both roles share one launch (heavy workgroups first).
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

namespace {

constexpr int IB = 4;  // batch entries per lane
constexpr int PF = 8;  // point loads in flight per lane in the adjoint

template <int EB>
struct Elem;
template <>
struct Elem<1> { using T = uint8_t; };
template <>
struct Elem<2> { using T = uint16_t; };
template <>
struct Elem<4> { using T = uint32_t; };
template <>
struct Elem<8> { using T = uint2; };
template <>
struct Elem<16> { using T = uint4; };

// I: type of the work-item index and its div / mod (uint32_t when every
// item number fits, which avoids the 64-bit division sequence per item)
template <typename E, typename I>
__global__ __launch_bounds__(256) void index_gather_kernel(
    const E* __restrict__ x, const int64_t* __restrict__ idx, E* __restrict__ out,
    int64_t nbatch, int64_t ngrid, int64_t npts) {
  const I nchunk = (I)((nbatch + IB - 1) / IB);
  const I items = (I)npts * nchunk;
  const I np = (I)npts;
  const I stride = (I)gridDim.x * blockDim.x;
  for (I t = (I)blockIdx.x * blockDim.x + threadIdx.x; t < items; t += stride) {
    const int64_t p = (int64_t)(t % np);
    const int64_t n0 = (int64_t)(t / np) * IB;
    const int64_t c = idx[p];
    const bool ok = c >= 0 && c < ngrid;
    E v[IB];
#pragma unroll
    for (int j = 0; j < IB; ++j) {
      v[j] = E{};
      if (ok && n0 + j < nbatch) v[j] = x[(n0 + j) * ngrid + c];
    }
#pragma unroll
    for (int j = 0; j < IB; ++j)
      if (n0 + j < nbatch) out[(n0 + j) * npts + p] = v[j];
  }
}

template <typename S, int L>
struct Acc {
  S v[L];
};

// Adds the points order[k0..k1) of batch entries n0..n0+IB-1 to acc in
// ascending k, PF loads in flight at a time.
template <typename S, int L>
__device__ __forceinline__ void sum_run(const S* __restrict__ x,
                                        const int64_t* __restrict__ order,
                                        int64_t k0, int64_t k1, int64_t n0,
                                        int64_t nbatch, int64_t npts,
                                        Acc<S, L> (&acc)[IB]) {
  for (int64_t k = k0; k < k1; k += PF) {
    int64_t q[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      q[u] = k + u < k1 ? order[k + u] : -1;
      if (q[u] >= npts) q[u] = -1;
    }
#pragma unroll
    for (int j = 0; j < IB; ++j) {
      if (n0 + j >= nbatch) break;
      S v[PF][L];
#pragma unroll
      for (int u = 0; u < PF; ++u)
#pragma unroll
        for (int l = 0; l < L; ++l)
          v[u][l] = q[u] >= 0 ? x[((n0 + j) * npts + q[u]) * L + l] : S(0);
#pragma unroll
      for (int u = 0; u < PF; ++u)
        if (q[u] >= 0)
#pragma unroll
          for (int l = 0; l < L; ++l) acc[j].v[l] += v[u][l];
    }
  }
}

template <typename S, int L>
__device__ __forceinline__ void zero_acc(Acc<S, L> (&acc)[IB]) {
#pragma unroll
  for (int j = 0; j < IB; ++j)
#pragma unroll
    for (int l = 0; l < L; ++l) acc[j].v[l] = S(0);
}

// Light cells: a lane per (cell, batch chunk), grid-stride over the
// workgroups vb of vg given to this role.
template <typename S, int L, typename I>
__device__ __forceinline__ void adjoint_light(
    const S* __restrict__ x, const int64_t* __restrict__ order,
    const int64_t* __restrict__ starts, S* __restrict__ out, int64_t nbatch,
    int64_t ngrid, int64_t npts, int64_t heavy_min, I vb, I vg) {
  const I nchunk = (I)((nbatch + IB - 1) / IB);
  const I items = (I)ngrid * nchunk;
  const I ng = (I)ngrid;
  const I stride = vg * (I)blockDim.x;
  for (I t = vb * (I)blockDim.x + threadIdx.x; t < items; t += stride) {
    const int64_t c = (int64_t)(t % ng);
    const int64_t n0 = (int64_t)(t / ng) * IB;
    int64_t k0 = starts[c], k1 = starts[c + 1];
    if (heavy_min > 0 && k1 - k0 > heavy_min) continue;  // a heavy cell
    k0 = k0 < 0 ? 0 : k0;
    k1 = k1 > npts ? npts : k1;
    Acc<S, L> acc[IB];
    zero_acc(acc);
    sum_run<S, L>(x, order, k0, k1, n0, nbatch, npts, acc);
#pragma unroll
    for (int j = 0; j < IB; ++j) {
      if (n0 + j >= nbatch) break;
      S* dst = out + ((n0 + j) * ngrid + c) * L;
#pragma unroll
      for (int l = 0; l < L; ++l) dst[l] = acc[j].v[l];
    }
  }
}

// Heavy cells: one workgroup per (heavy cell, batch chunk), over the
// workgroups vb of vg given to this role: thread i sums the i-th of 256
// contiguous slices of the cell's run in order, then a fixed pairwise LDS
// tree adds the 256 partial sums.
template <typename S, int L>
__device__ __forceinline__ void adjoint_heavy(
    const S* __restrict__ x, const int64_t* __restrict__ order,
    const int64_t* __restrict__ starts, const int64_t* __restrict__ heavy,
    int64_t nheavy, S* __restrict__ out, int64_t nbatch, int64_t ngrid,
    int64_t npts, int64_t vb, int64_t vg, S (*red)[IB * L]) {
  const int tid = threadIdx.x;
  const int64_t nchunk = (nbatch + IB - 1) / IB;
  for (int64_t b = vb; b < nheavy * nchunk; b += vg) {
    const int64_t c = heavy[b % nheavy];
    const int64_t n0 = (b / nheavy) * IB;
    const bool ok = c >= 0 && c < ngrid;
    int64_t k0 = ok ? starts[c] : 0, k1 = ok ? starts[c + 1] : 0;
    k0 = k0 < 0 ? 0 : k0;
    k1 = k1 > npts ? npts : k1;
    const int64_t per = k1 > k0 ? (k1 - k0 + 255) / 256 : 0;
    int64_t a = k0 + tid * per, e = a + per;
    a = a > k1 ? k1 : a;
    e = e > k1 ? k1 : e;
    Acc<S, L> acc[IB];
    zero_acc(acc);
    sum_run<S, L>(x, order, a, e, n0, nbatch, npts, acc);
#pragma unroll
    for (int j = 0; j < IB; ++j)
#pragma unroll
      for (int l = 0; l < L; ++l) red[tid][j * L + l] = acc[j].v[l];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (tid < h)
#pragma unroll
        for (int i = 0; i < IB * L; ++i) red[tid][i] += red[tid + h][i];
      __syncthreads();
    }
    if (ok && tid < IB * L) {
      const int j = tid / L, l = tid % L;
      if (n0 + j < nbatch) out[((n0 + j) * ngrid + c) * L + l] = red[0][tid];
    }
    __syncthreads();
  }
}

// One launch for both roles: the first hb workgroups take the heavy cells
// (so they start first), the rest the light ones; the two write disjoint
// cells and run side by side.
template <typename S, int L, typename I>
__global__ __launch_bounds__(256) void index_adjoint_kernel(
    const S* __restrict__ x, const int64_t* __restrict__ order,
    const int64_t* __restrict__ starts, const int64_t* __restrict__ heavy,
    int64_t nheavy, S* __restrict__ out, int64_t nbatch, int64_t ngrid,
    int64_t npts, int64_t heavy_min, unsigned hb) {
  __shared__ S red[256][IB * L];
  if (blockIdx.x < hb)
    adjoint_heavy<S, L>(x, order, starts, heavy, nheavy, out, nbatch, ngrid, npts,
                        (int64_t)blockIdx.x, (int64_t)hb, red);
  else
    adjoint_light<S, L, I>(x, order, starts, out, nbatch, ngrid, npts, heavy_min,
                           (I)(blockIdx.x - hb), (I)(gridDim.x - hb));
}

unsigned launch_blocks(int64_t items) {
  // grid-stride: enough workgroups to fill 256 CUs several times over
  const int64_t want = (items + 255) / 256;
  const int64_t cap = 256 * 64;
  return (unsigned)(want < 1 ? 1 : (want > cap ? cap : want));
}

template <int EB>
int launch_gather(const void* x, const int64_t* idx, void* out, int64_t nbatch,
                  int64_t ngrid, int64_t npts, hipStream_t s) {
  using E = typename Elem<EB>::T;
  const int64_t items = npts * ((nbatch + IB - 1) / IB);
  if (items + 256LL * 64 * 256 < (1LL << 32))
    index_gather_kernel<E, uint32_t><<<launch_blocks(items), 256, 0, s>>>(
        (const E*)x, idx, (E*)out, nbatch, ngrid, npts);
  else
    index_gather_kernel<E, int64_t><<<launch_blocks(items), 256, 0, s>>>(
        (const E*)x, idx, (E*)out, nbatch, ngrid, npts);
  TNL_LAUNCH_CHECK("index_gather");
  return TNL_OK;
}

template <typename S, int L>
int launch_adjoint(const void* x, const int64_t* order, const int64_t* starts,
                   const int64_t* heavy, int64_t nheavy, int64_t heavy_min,
                   void* out, int64_t nbatch, int64_t ngrid, int64_t npts,
                   hipStream_t s) {
  const int64_t nchunk = (nbatch + IB - 1) / IB;
  const int64_t hm = nheavy > 0 ? heavy_min : 0;
  const int64_t nh = nheavy * nchunk;
  const unsigned hb = (unsigned)(nh < 65536 ? nh : 65536);  // heavy-role workgroups
  const unsigned grid = hb + launch_blocks(ngrid * nchunk);
  if (ngrid * nchunk + 256LL * 64 * 256 < (1LL << 32))
    index_adjoint_kernel<S, L, uint32_t><<<grid, 256, 0, s>>>(
        (const S*)x, order, starts, heavy, nheavy, (S*)out, nbatch, ngrid, npts, hm, hb);
  else
    index_adjoint_kernel<S, L, int64_t><<<grid, 256, 0, s>>>(
        (const S*)x, order, starts, heavy, nheavy, (S*)out, nbatch, ngrid, npts, hm, hb);
  TNL_LAUNCH_CHECK("index_adjoint");
  return TNL_OK;
}

}  // namespace

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_index(const void* x, const int64_t* idx, void* out,
                         int elem_bytes, int64_t nbatch, int64_t ngrid,
                         int64_t npts, void* stream) {
  if (nbatch < 0 || ngrid < 0 || npts < 0)
    return fail(TNL_EINVAL, "nbatch, ngrid and npts must be >= 0");
  if (nbatch == 0 || npts == 0) return TNL_OK;
  if (ngrid == 0) return fail(TNL_EINVAL, "cannot index an empty grid");
  if (!x || !idx || !out) return fail(TNL_EINVAL, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  switch (elem_bytes) {
    case 1: return launch_gather<1>(x, idx, out, nbatch, ngrid, npts, s);
    case 2: return launch_gather<2>(x, idx, out, nbatch, ngrid, npts, s);
    case 4: return launch_gather<4>(x, idx, out, nbatch, ngrid, npts, s);
    case 8: return launch_gather<8>(x, idx, out, nbatch, ngrid, npts, s);
    case 16: return launch_gather<16>(x, idx, out, nbatch, ngrid, npts, s);
    default: return fail(TNL_EINVAL, "elem_bytes must be 1, 2, 4, 8 or 16");
  }
}

extern "C" int tnl_index_adjoint(const void* x, const int64_t* order,
                                 const int64_t* starts, const int64_t* heavy,
                                 int64_t nheavy, int64_t heavy_min, void* out,
                                 int dtype, int is_complex, int64_t nbatch,
                                 int64_t ngrid, int64_t npts, void* stream) {
  if (nbatch < 0 || ngrid < 0 || npts < 0 || nheavy < 0)
    return fail(TNL_EINVAL, "nbatch, ngrid, npts and nheavy must be >= 0");
  if (nheavy > 0 && (heavy_min < 1 || !heavy))
    return fail(TNL_EINVAL, "a heavy-cell list needs heavy_min >= 1 and a pointer");
  const bool integer = dtype == TNL_I8 || dtype == TNL_U8 || dtype == TNL_I16 ||
                       dtype == TNL_I32 || dtype == TNL_I64;
  if (dtype != TNL_F32 && dtype != TNL_F64 && !integer)
    return fail(TNL_EINVAL, "index_adjoint dtype must be float32, float64 or an integer type");
  if (integer && is_complex) return fail(TNL_EINVAL, "integer index_adjoint has no complex form");
  if (nbatch == 0 || ngrid == 0) return TNL_OK;
  if (!out || !starts || (npts > 0 && (!x || !order)))
    return fail(TNL_EINVAL, "null pointer");
  hipStream_t s = (hipStream_t)stream;
  // integer sums are exact and order-free, wrapping like the sequential
  // index_add_ of the same type; they run in the unsigned type of the same
  // width, whose addition is the two's-complement sum bit for bit (signed
  // overflow would be undefined behaviour)
  switch (dtype) {
    case TNL_I8:
    case TNL_U8:
      return launch_adjoint<uint8_t, 1>(x, order, starts, heavy, nheavy, heavy_min, out, nbatch,
                                        ngrid, npts, s);
    case TNL_I16:
      return launch_adjoint<uint16_t, 1>(x, order, starts, heavy, nheavy, heavy_min, out,
                                         nbatch, ngrid, npts, s);
    case TNL_I32:
      return launch_adjoint<uint32_t, 1>(x, order, starts, heavy, nheavy, heavy_min, out,
                                         nbatch, ngrid, npts, s);
    case TNL_I64:
      return launch_adjoint<uint64_t, 1>(x, order, starts, heavy, nheavy, heavy_min, out,
                                         nbatch, ngrid, npts, s);
    default:
      break;
  }
  if (dtype == TNL_F32)
    return is_complex ? launch_adjoint<float, 2>(x, order, starts, heavy, nheavy, heavy_min,
                                                 out, nbatch, ngrid, npts, s)
                      : launch_adjoint<float, 1>(x, order, starts, heavy, nheavy, heavy_min,
                                                 out, nbatch, ngrid, npts, s);
  return is_complex ? launch_adjoint<double, 2>(x, order, starts, heavy, nheavy, heavy_min,
                                                out, nbatch, ngrid, npts, s)
                    : launch_adjoint<double, 1>(x, order, starts, heavy, nheavy, heavy_min,
                                                out, nbatch, ngrid, npts, s);
}
