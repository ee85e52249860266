// Plan-driven interpolation (ungrid) and gridding kernels, 2-D and 3-D,
// norm 1 (separable weights).
//
// ungrid: one workgroup per (subproblem, channel chunk).  A subproblem is
//   <= max_sub points of one bin.  The workgroup stages the bin's grid tile
//   plus a halo of ntap-1 cells per axis (wrapped, or zero outside the grid
//   under zero padding) for CC channels in LDS, then each lane gathers the
//   taps of one point from LDS for all CC channels at once: the per-point
//   weights are computed once and reused for every channel.
// grid: "owner computes".  One workgroup per (work item, channel chunk); a
//   work item is a gridding tile with the list of points touching it.  Taps
//   inside the tile are accumulated with LDS float atomics, then the tile is
//   written to HBM once with coalesced stores.  No zero-fill pass and no
//   global atomics, except for tiles with more than g_cap touches, which
//   are split across workgroups and combined with global float atomics into
//   a region zeroed by a small pre-pass.
// Workgroup ids are remapped so each XCD walks a contiguous range of tiles
// (neighbouring halos then share that XCD's L2).
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

constexpr int BIN_THREADS = 256;
constexpr int NUM_XCD = 8;
constexpr int64_t LDS_BUDGET = 64 * 1024;

struct BinGeom {
  int tile[3];
  int ntiles[3];
  int win[3];  // tile + ntap - 1 (ungrid window)
  int wcells;
  int tcells;
  int64_t npts;
  int64_t nbins;
};

template <bool CPLX>
struct Ch;
template <>
struct Ch<false> {
  using T = float;
  __device__ static T zero() { return 0.f; }
  __device__ static void fma(T& a, float w, T v) { a = fmaf(w, v, a); }
};
template <>
struct Ch<true> {
  using T = float2;
  __device__ static T zero() { return make_float2(0.f, 0.f); }
  __device__ static void fma(T& a, float w, T v) {
    a.x = fmaf(w, v.x, a.x);
    a.y = fmaf(w, v.y, a.y);
  }
};

__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nwg) {
  const int64_t per = nwg / NUM_XCD;  // nwg is a multiple of NUM_XCD
  return (bid % NUM_XCD) * per + bid / NUM_XCD;
}

template <int NDIM, bool CPLX, int CC>
__global__ __launch_bounds__(BIN_THREADS) void ungrid_binned_kernel(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    InterpGeom G, BinGeom B, const int* __restrict__ perm,
    const float* __restrict__ slocs, const int4* __restrict__ sub,
    const int* __restrict__ nsub_ptr) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const int64_t s = xcd_remap(blockIdx.x, gridDim.x);
  if (s >= *nsub_ptr) return;
  const int4 sp = sub[s];
  const int bin = sp.x, start = sp.y, count = sp.z;
  const int64_t c0 = (int64_t)blockIdx.y * CC;
  int nc = (int)(nbatch - c0);
  nc = nc < CC ? nc : CC;

  // tile origin per (padded) axis
  int origin[3];
  {
    int r = bin;
    for (int d = 2; d >= 0; --d) {
      const int ti = r % B.ntiles[d];
      r /= B.ntiles[d];
      origin[d] = ti * B.tile[d];
    }
  }
  // ---- stage window: [cell][channel] in LDS, coalesced along z in HBM
  const T* __restrict__ v = reinterpret_cast<const T*>(vals);
  const int w12 = B.win[1] * B.win[2];
  const int total = B.wcells * CC;
  for (int idx = threadIdx.x; idx < total; idx += BIN_THREADS) {
    const int c = idx / B.wcells;
    const int cell = idx - c * B.wcells;
    const int wx = cell / w12;
    const int rem = cell - wx * w12;
    const int wy = rem / B.win[2];
    const int wz = rem - wy * B.win[2];
    int64_t g[3] = {origin[0] + wx, origin[1] + wy, origin[2] + wz};
    bool ok = c < nc;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (G.pad == TNL_PAD_CIRCULAR) {
        if (g[d] >= G.n[d]) g[d] = mod_pos(g[d], G.n[d]);
      } else {
        ok = ok && g[d] < G.n[d];
      }
    }
    T val = C::zero();
    if (ok)
      val = v[(c0 + c) * G.ncells + (g[0] * G.n[1] + g[1]) * G.n[2] + g[2]];
    lds[cell * CC + c] = val;
  }
  __syncthreads();

  // ---- gather: one point per lane
  for (int i = threadIdx.x; i < count; i += BIN_THREADS) {
    const int q = start + i;
    const int p = perm[q];
    float w[3][8];
    int lo[3], cnt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d < 3 - NDIM) {
        lo[d] = 0;
        cnt[d] = 1;
        w[d][0] = 1.0f;
        continue;
      }
      const float t = slocs[(int64_t)q * NDIM + d - (3 - NDIM)];
      const TapRange r = tap_range(t, G.half[d], G.ntap[d], G.n[d], G.pad);
      cnt[d] = r.count;
      const int64_t gw =
          G.pad == TNL_PAD_CIRCULAR ? mod_pos(r.g0, G.n[d]) : r.g0;
      lo[d] = (int)(gw - origin[d]);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        w[d][k] = 0.0f;
        if (k < r.count) {
          const float x =
              fminf(fabsf((float)(r.g0 + k) - t) * G.inv_half[d], 1.0f);
          w[d][k] = window(G.kernel, x, G.beta);
        }
      }
    }
    T acc[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c) acc[c] = C::zero();
#pragma unroll
    for (int kx = 0; kx < (NDIM >= 3 ? 8 : 1); ++kx) {
      if (kx >= cnt[0]) break;
#pragma unroll
      for (int ky = 0; ky < 8; ++ky) {
        if (ky >= cnt[1]) break;
        const float wxy = w[0][kx] * w[1][ky];
        const int row = ((lo[0] + kx) * B.win[1] + (lo[1] + ky)) * B.win[2] +
                        lo[2];
#pragma unroll
        for (int kz = 0; kz < 8; ++kz) {
          if (kz >= cnt[2]) break;
          const float wt = wxy * w[2][kz];
          const T* cellp = lds + (row + kz) * CC;
#pragma unroll
          for (int c = 0; c < CC; ++c) C::fma(acc[c], wt, cellp[c]);
        }
      }
    }
    T* __restrict__ o = reinterpret_cast<T*>(out);
#pragma unroll
    for (int c = 0; c < CC; ++c)
      if (c < nc) o[(c0 + c) * B.npts + p] = acc[c];
  }
}

// Pre-pass for gridding: zero the output region of every split (shared)
// tile, once per tile (by its first item).
template <bool CPLX>
__global__ void grid_zero_shared_kernel(float* __restrict__ out,
                                        int64_t nbatch, InterpGeom G,
                                        BinGeom B,
                                        const int4* __restrict__ items,
                                        const int* __restrict__ nitems_ptr,
                                        const int* __restrict__ gstarts) {
  const int nitems = *nitems_ptr;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int4 item = items[it];
    if (!item.w || item.y != gstarts[item.x]) continue;
    int origin[3];
    int r = item.x;
    for (int d = 2; d >= 0; --d) {
      origin[d] = (r % B.ntiles[d]) * B.tile[d];
      r /= B.ntiles[d];
    }
    const int comp = CPLX ? 2 : 1;
    const int t12 = B.tile[1] * B.tile[2];
    const int64_t total = (int64_t)nbatch * B.tcells * comp;
    for (int64_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int64_t c = idx / (B.tcells * comp);
      int e = (int)(idx - c * B.tcells * comp);
      const int part = e % comp;
      const int cell = e / comp;
      const int lx = cell / t12, rem = cell - lx * t12;
      const int ly = rem / B.tile[2], lz = rem - ly * B.tile[2];
      const int64_t gx = origin[0] + lx, gy = origin[1] + ly,
                    gz = origin[2] + lz;
      if (gx >= G.n[0] || gy >= G.n[1] || gz >= G.n[2]) continue;
      out[(c * G.ncells + (gx * G.n[1] + gy) * G.n[2] + gz) * comp + part] =
          0.0f;
    }
  }
}

template <int NDIM, bool CPLX, int CC>
__global__ __launch_bounds__(BIN_THREADS) void grid_binned_kernel(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    InterpGeom G, BinGeom B, const int* __restrict__ perm,
    const float* __restrict__ slocs, const int* __restrict__ tlist,
    const int4* __restrict__ items, const int* __restrict__ nitems_ptr) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  constexpr int COMP = CPLX ? 2 : 1;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  float* acc = reinterpret_cast<float*>(lds_raw);  // [cell][CC][COMP]
  const int64_t it = xcd_remap(blockIdx.x, gridDim.x);
  if (it >= *nitems_ptr) return;
  const int4 item = items[it];
  const int tile = item.x, start = item.y, count = item.z;
  const bool shared = item.w != 0;
  const int64_t c0 = (int64_t)blockIdx.y * CC;
  int nc = (int)(nbatch - c0);
  nc = nc < CC ? nc : CC;
  int origin[3];
  {
    int r = tile;
    for (int d = 2; d >= 0; --d) {
      origin[d] = (r % B.ntiles[d]) * B.tile[d];
      r /= B.ntiles[d];
    }
  }
  const int nacc = B.tcells * CC * COMP;
  for (int i = threadIdx.x; i < nacc; i += BIN_THREADS) acc[i] = 0.0f;
  __syncthreads();

  const T* __restrict__ v = reinterpret_cast<const T*>(vals);
  for (int i = threadIdx.x; i < count; i += BIN_THREADS) {
    const int q = tlist[start + i];
    const int p = perm[q];
    // per axis: weight and local (in-tile) coordinate of every tap; taps
    // outside this tile get l = -1.  Under circular padding a footprint
    // wider than the grid wraps, so in-tile taps need not be contiguous.
    float w[3][8];
    int l[3][8];
    int cnt[3];
    bool live = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d < 3 - NDIM) {
        cnt[d] = 1;
        w[d][0] = 1.0f;
        l[d][0] = 0;
        continue;
      }
      const float t = slocs[(int64_t)q * NDIM + d - (3 - NDIM)];
      const TapRange r = tap_range(t, G.half[d], G.ntap[d], G.n[d], G.pad);
      cnt[d] = r.count;
      bool any = false;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        w[d][k] = 0.0f;
        l[d][k] = -1;
        if (k < r.count) {
          const int64_t g = G.pad == TNL_PAD_CIRCULAR
                                ? mod_pos(r.g0 + k, G.n[d])
                                : r.g0 + k;
          const int64_t lc = g - origin[d];
          if (lc >= 0 && lc < B.tile[d]) {
            const float x =
                fminf(fabsf((float)(r.g0 + k) - t) * G.inv_half[d], 1.0f);
            w[d][k] = window(G.kernel, x, G.beta);
            l[d][k] = (int)lc;
            any = true;
          }
        }
      }
      live = live && any;
    }
    if (!live) continue;
    T val[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c)
      val[c] = c < nc ? v[(c0 + c) * B.npts + p] : C::zero();
#pragma unroll
    for (int kx = 0; kx < (NDIM >= 3 ? 8 : 1); ++kx) {
      if (kx >= cnt[0]) break;
      if (l[0][kx] < 0) continue;
#pragma unroll
      for (int ky = 0; ky < 8; ++ky) {
        if (ky >= cnt[1]) break;
        if (l[1][ky] < 0) continue;
        const float wxy = w[0][kx] * w[1][ky];
        const int row = (l[0][kx] * B.tile[1] + l[1][ky]) * B.tile[2];
#pragma unroll
        for (int kz = 0; kz < 8; ++kz) {
          if (kz >= cnt[2]) break;
          if (l[2][kz] < 0) continue;
          const float wt = wxy * w[2][kz];
          float* cellp = acc + (row + l[2][kz]) * (CC * COMP);
#pragma unroll
          for (int c = 0; c < CC; ++c) {
            if constexpr (CPLX) {
              atomicAdd(cellp + 2 * c, wt * val[c].x);
              atomicAdd(cellp + 2 * c + 1, wt * val[c].y);
            } else {
              atomicAdd(cellp + c, wt * val[c]);
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- write the tile: coalesced along z, one channel plane at a time
  const int t12 = B.tile[1] * B.tile[2];
  T* __restrict__ o = reinterpret_cast<T*>(out);
  const int total = B.tcells * nc;
  for (int idx = threadIdx.x; idx < total; idx += BIN_THREADS) {
    const int c = idx / B.tcells;
    const int cell = idx - c * B.tcells;
    const int lx = cell / t12, rem = cell - lx * t12;
    const int ly = rem / B.tile[2], lz = rem - ly * B.tile[2];
    const int64_t gx = origin[0] + lx, gy = origin[1] + ly, gz = origin[2] + lz;
    if (gx >= G.n[0] || gy >= G.n[1] || gz >= G.n[2]) continue;
    const int64_t gi = (c0 + c) * G.ncells + (gx * G.n[1] + gy) * G.n[2] + gz;
    const float* a = acc + cell * (CC * COMP) + c * COMP;
    if (!shared) {
      if constexpr (CPLX)
        o[gi] = make_float2(a[0], a[1]);
      else
        o[gi] = a[0];
    } else {
      if constexpr (CPLX) {
        unsafeAtomicAdd(out + 2 * gi, a[0]);
        unsafeAtomicAdd(out + 2 * gi + 1, a[1]);
      } else {
        unsafeAtomicAdd(out + gi, a[0]);
      }
    }
  }
}

static BinGeom make_bin_geom(const tnl_plan_layout& L, bool adjoint) {
  BinGeom B;
  std::memset(&B, 0, sizeof(B));
  B.wcells = 1;
  B.tcells = 1;
  for (int d = 0; d < 3; ++d) {
    B.tile[d] = adjoint ? L.g_tile[d] : L.u_tile[d];
    B.ntiles[d] = adjoint ? L.g_ntiles[d] : L.u_ntiles[d];
    B.win[d] = L.u_tile[d] + L.ntaps[d] - 1;
    B.wcells *= B.win[d];
    B.tcells *= B.tile[d];
  }
  B.npts = L.npts;
  B.nbins = L.u_nbins;
  return B;
}

bool binned_supported(const InterpGeom& G, const tnl_plan_layout& L) {
  if (G.ndim < 2 || G.norm != 1) return false;
  for (int d = 0; d < 3; ++d)
    if (G.ntap[d] > 8 || G.ntap[d] != L.ntaps[d]) return false;
  const BinGeom Bu = make_bin_geom(L, false);
  const BinGeom Bg = make_bin_geom(L, true);
  // at least one channel must fit in LDS for both directions
  return (int64_t)Bu.wcells * 8 <= LDS_BUDGET &&
         (int64_t)Bg.tcells * 8 <= LDS_BUDGET;
}

static int pick_cc(int64_t nbatch, int64_t cells, int comp_bytes) {
  int cc = 1;
  while (cc < 8 && cc < nbatch && (int64_t)cells * comp_bytes * cc * 2 <= LDS_BUDGET)
    cc *= 2;
  return cc;
}

template <int NDIM, bool CPLX, int CC>
static int launch_binned(bool adjoint, const float* vals, float* out,
                         int64_t nbatch, const InterpGeom& G,
                         const tnl_plan_layout& L, const char* ws,
                         hipStream_t s) {
  const BinGeom B = make_bin_geom(L, adjoint);
  const int comp_bytes = CPLX ? 8 : 4;
  const unsigned nchunk = (unsigned)((nbatch + CC - 1) / CC);
  const int* perm = (const int*)(ws + L.off_perm);
  const float* slocs = (const float*)(ws + L.off_slocs);
  if (!adjoint) {
    const int64_t nwg = (L.sub_capacity + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const size_t lds = (size_t)B.wcells * CC * comp_bytes;
    ungrid_binned_kernel<NDIM, CPLX, CC>
        <<<dim3((unsigned)nwg, nchunk), BIN_THREADS, lds, s>>>(
            vals, out, nbatch, G, B, perm, slocs,
            (const int4*)(ws + L.off_sub),
            (const int*)(ws + L.off_suboff) + L.u_nbins);
    TNL_LAUNCH_CHECK("ungrid_binned");
  } else {
    const int* nitems = (const int*)(ws + L.off_itemoff) + L.g_ntiles_total;
    const int4* items = (const int4*)(ws + L.off_items);
    grid_zero_shared_kernel<CPLX><<<1024, 256, 0, s>>>(
        out, nbatch, G, B, items, nitems, (const int*)(ws + L.off_gstarts));
    TNL_LAUNCH_CHECK("grid_zero_shared");
    const int64_t nwg = (L.item_capacity + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const size_t lds = (size_t)B.tcells * CC * comp_bytes;
    grid_binned_kernel<NDIM, CPLX, CC>
        <<<dim3((unsigned)nwg, nchunk), BIN_THREADS, lds, s>>>(
            vals, out, nbatch, G, B, perm, slocs,
            (const int*)(ws + L.off_tlist), items, nitems);
    TNL_LAUNCH_CHECK("grid_binned");
  }
  return TNL_OK;
}

template <int NDIM, bool CPLX>
static int dispatch_cc(bool adjoint, const float* vals, float* out,
                       int64_t nbatch, const InterpGeom& G,
                       const tnl_plan_layout& L, const char* ws,
                       hipStream_t s) {
  const BinGeom B = make_bin_geom(L, adjoint);
  const int cc = pick_cc(nbatch, adjoint ? B.tcells : B.wcells, CPLX ? 8 : 4);
  switch (cc) {
    case 8: return launch_binned<NDIM, CPLX, 8>(adjoint, vals, out, nbatch, G, L, ws, s);
    case 4: return launch_binned<NDIM, CPLX, 4>(adjoint, vals, out, nbatch, G, L, ws, s);
    case 2: return launch_binned<NDIM, CPLX, 2>(adjoint, vals, out, nbatch, G, L, ws, s);
    default: return launch_binned<NDIM, CPLX, 1>(adjoint, vals, out, nbatch, G, L, ws, s);
  }
}

int interp_binned(bool adjoint, const float* vals, float* out, int64_t nbatch,
                  const InterpGeom& G, int is_complex, const tnl_plan_layout& L,
                  const void* plan, hipStream_t s) {
  const char* ws = (const char*)plan;
  if (G.ndim == 3)
    return is_complex ? dispatch_cc<3, true>(adjoint, vals, out, nbatch, G, L, ws, s)
                      : dispatch_cc<3, false>(adjoint, vals, out, nbatch, G, L, ws, s);
  return is_complex ? dispatch_cc<2, true>(adjoint, vals, out, nbatch, G, L, ws, s)
                    : dispatch_cc<2, false>(adjoint, vals, out, nbatch, G, L, ws, s);
}

}  // namespace tnl
