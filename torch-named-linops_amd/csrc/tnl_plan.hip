// Point-binning plan: a stable counting sort of the nonuniform points into
// spatial tiles, built entirely on the device (no host synchronisation, so
// it can be captured in a HIP graph).
//
// Two tilings share one workspace:
//   * interpolation bins (u_*): each point belongs to the bin holding its
//     first kept tap; ungrid stages bin + halo of the grid in LDS.
//   * gridding tiles (g_*): each tile lists every point whose taps touch it;
//     grid accumulates a tile in LDS and writes it once ("owner computes"),
//     so the output needs neither zero-fill nor global atomics except for
//     the few tiles dense enough to be split across workgroups.
#include <algorithm>

#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

constexpr int PLAN_THREADS = 256;

struct PlanDev {
  int ndim;
  int64_t n[3];
  float half[3];
  int ntap[3];
  int pad;
  int utile[3], untiles[3];
  int gtile[3], gntiles[3];
  int64_t u_nbins, g_ntiles_total;
  int64_t npts;
  int max_sub;
  int g_cap;
};

static int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

static int default_u_tile(int ndim) {
  return ndim == 3 ? 8 : ndim == 2 ? 32 : 1024;
}
static void default_g_tile(int ndim, int t[3]) {
  // innermost (contiguous) axis longest: coalesced tile writes
  if (ndim == 3) { t[0] = 8; t[1] = 8; t[2] = 8; }
  else if (ndim == 2) { t[0] = 1; t[1] = 16; t[2] = 32; }
  else { t[0] = 1; t[1] = 1; t[2] = 4096; }
}

int plan_layout(int64_t npts, int ndim, const int64_t* grid_size,
                const float* width, int pad_mode, const tnl_tuning* tune,
                tnl_plan_layout* L) {
  if (ndim < 1 || ndim > 3) return fail(TNL_EINVAL, "ndim must be 1, 2 or 3");
  if (npts < 0) return fail(TNL_EINVAL, "npts must be >= 0");
  if (npts > (int64_t(1) << 28))
    return fail(TNL_EINVAL, "plan supports at most 2^28 points");
  if (pad_mode != TNL_PAD_CIRCULAR && pad_mode != TNL_PAD_ZERO)
    return fail(TNL_EINVAL, "Unrecognized pad_mode");
  std::memset(L, 0, sizeof(*L));
  L->ndim = ndim;
  L->npts = npts;
  L->pad_mode = pad_mode;
  int gt[3];
  default_g_tile(ndim, gt);
  int ut_def[3] = {1, 1, 1};
  for (int d = 3 - ndim; d < 3; ++d) ut_def[d] = default_u_tile(ndim);
  if (tune) {
    for (int d = 0; d < 3; ++d) {
      if (tune->u_tile[d] > 0) ut_def[d] = tune->u_tile[d];
      if (tune->g_tile[d] > 0) gt[d] = tune->g_tile[d];
    }
  }
  int64_t n[3] = {1, 1, 1};
  int64_t touch_per_pt = 1;
  for (int d = 0; d < 3; ++d) {
    L->grid_size[d] = 1;
    L->width[d] = 1.0f;
    L->ntaps[d] = 1;
    L->u_tile[d] = 1;
    L->u_ntiles[d] = 1;
    L->g_tile[d] = 1;
    L->g_ntiles[d] = 1;
  }
  for (int d = 0; d < ndim; ++d) {
    const int gd = 3 - ndim + d;
    if (grid_size[d] < 1) return fail(TNL_EINVAL, "grid_size must be >= 1");
    if (!(width[d] > 0.0f) || !std::isfinite(width[d]))
      return fail(TNL_EINVAL, "kernel width must be positive and finite");
    n[gd] = grid_size[d];
    L->grid_size[gd] = grid_size[d];
    L->width[gd] = width[d];
    const int ntap = (int)std::ceil((double)width[d]) + 1;
    if (ntap > TNL_MAX_TAPS) return fail(TNL_EINVAL, "kernel width too large");
    L->ntaps[gd] = ntap;
    int64_t ut = ut_def[gd];
    if (ut > n[gd]) ut = n[gd];
    L->u_tile[gd] = (int32_t)ut;
    L->u_ntiles[gd] = (int32_t)((n[gd] + ut - 1) / ut);
    int64_t g = std::max<int64_t>(gt[gd], ntap);
    if (2 * g >= n[gd]) g = n[gd];  // one tile along this axis
    L->g_tile[gd] = (int32_t)g;
    L->g_ntiles[gd] = (int32_t)((n[gd] + g - 1) / g);
    const int per = L->g_ntiles[gd] == 1 ? 1 : (n[gd] % g == 0 ? 2 : 3);
    touch_per_pt *= per;
  }
  L->u_nbins = (int64_t)L->u_ntiles[0] * L->u_ntiles[1] * L->u_ntiles[2];
  L->g_ntiles_total = (int64_t)L->g_ntiles[0] * L->g_ntiles[1] * L->g_ntiles[2];
  L->max_sub = (tune && tune->max_sub > 0) ? tune->max_sub : 512;
  L->sub_capacity = L->u_nbins + (npts + L->max_sub - 1) / L->max_sub + 1;
  // touches per gridding work item: 4096 in 3-D (the 3-D radial workload,
  // 8 x 256^3, 20000 x 512 points, grids in 1.82 ms against 1.91 ms with
  // 2048: fewer split tiles and atomic adds; 6144 and up are slower, as
  // items past 4096 touches keep their tile order instead of the block
  // mask order), 2048 in 2-D
  L->g_cap = (tune && tune->g_cap > 0) ? tune->g_cap : (ndim == 3 ? 4096 : 2048);
  L->lds_budget = (tune && tune->lds_budget > 0) ? tune->lds_budget : 53248;
  if (L->lds_budget > 160 * 1024)
    return fail(TNL_EINVAL, "lds_budget above the 160 KiB LDS of a CU");
  {
    const int ga = tune ? tune->grid_accum : TNL_GRID_ACCUM_AUTO;
    if (ga < TNL_GRID_ACCUM_AUTO || ga > TNL_GRID_ACCUM_MFMA)
      return fail(TNL_EINVAL, "unknown grid_accum");
    const bool mf_tiles =
        (ndim == 3 && L->g_tile[0] == 8 && L->g_tile[1] == 8 && L->g_tile[2] == 8) ||
        (ndim == 2 && L->g_tile[1] == 16 && L->g_tile[2] == 32);
    if (ga == TNL_GRID_ACCUM_MFMA && !mf_tiles)
      return fail(TNL_EINVAL,
                  "the MFMA gridding accumulator needs 8x8x8 tiles (3-D) or 16x32 tiles (2-D)");
    L->grid_accum = ga == TNL_GRID_ACCUM_AUTO
                        ? (mf_tiles ? TNL_GRID_ACCUM_MFMA : TNL_GRID_ACCUM_LDS)
                        : ga;
  }
  L->touch_capacity = npts * touch_per_pt;
  L->item_capacity =
      L->g_ntiles_total + (L->touch_capacity + L->g_cap - 1) / L->g_cap + 1;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const int64_t nb1 = L->u_nbins + 1;  // + dead bin
  L->off_counts = take(4 * nb1);
  L->off_starts = take(4 * (nb1 + 1));
  L->off_cursor = take(4 * nb1);
  L->off_perm = take(4 * std::max<int64_t>(npts, 1));
  L->off_slocs = take(4 * std::max<int64_t>(npts, 1) * ndim);
  L->off_suboff = take(4 * (L->u_nbins + 2));
  L->off_sub = take(16 * L->sub_capacity);
  L->off_gcounts = take(4 * L->g_ntiles_total);
  L->off_gstarts = take(4 * (L->g_ntiles_total + 1));
  L->off_gcursor = take(4 * L->g_ntiles_total);
  L->off_tlist = take(4 * std::max<int64_t>(L->touch_capacity, 1));
  L->off_itemoff = take(4 * (L->g_ntiles_total + 1));
  L->off_items = take(16 * L->item_capacity);
  L->off_items_tmp = take(16 * L->item_capacity);
  // touch pairs of the one-wave matrix-core gridding kernel (4 spare pairs:
  // a wave reads its last touches 4 at a time)
  bool wave = ndim == 3 && L->grid_accum == TNL_GRID_ACCUM_MFMA;
  for (int d = 0; wave && d < 3; ++d)
    wave = L->g_tile[d] == 8 && L->ntaps[d] <= 5 && n[d] >= 16 && n[d] <= 1024;
  L->off_tpm = wave ? take(8 * (L->touch_capacity + 4)) : 0;
  L->off_wrec = wave ? take(64 * std::max<int64_t>(npts, 1)) : 0;
  L->wrec_kernel = -1;
  L->wrec_beta = 0.0f;
  L->bytes = off;
  return TNL_OK;
}

// ------------------------------------------------------------ device side
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// atomicAdd(&ctr[key], 1) for every active lane, merged per distinct key so
// a wave issues one atomic per key; slots follow lane order, which keeps the
// sort stable inside a wave.
__device__ __forceinline__ int agg_inc(int* ctr, int key, bool active) {
  const int lane = lane_id();
  uint64_t pending = __ballot(active);
  int out = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lkey = __shfl(key, leader);
    const bool mine = ((pending >> lane) & 1ull) && key == lkey;
    const uint64_t same = __ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(ctr + lkey, (int)__popcll(same));
    base = __shfl(base, leader);
    if (mine) out = base + (int)__popcll(same & ((1ull << lane) - 1ull));
    pending &= ~same;
  }
  return out;
}

// Bin of a point for interpolation, or P.u_nbins when it has no taps.
__device__ __forceinline__ int u_bin_of(const float* t, const PlanDev& P) {
  int64_t bin = 0;
  for (int d = 0; d < 3; ++d) {
    int64_t c = 0;
    if (d >= 3 - P.ndim) {
      const TapRange r = tap_range(t[d - (3 - P.ndim)], P.half[d], P.ntap[d],
                                   P.n[d], P.pad);
      if (r.count == 0) return (int)P.u_nbins;
      c = P.pad == TNL_PAD_CIRCULAR ? mod_pos(r.g0, P.n[d]) : r.g0;
    }
    bin = bin * P.untiles[d] + c / P.utile[d];
  }
  return (int)bin;
}

// Distinct gridding tiles touched along each axis (up to 3 per axis).
struct Touch {
  int cnt[3];
  int tile[3][3];
};

__device__ __forceinline__ bool touches_of(const float* t, const PlanDev& P,
                                           Touch& T) {
  for (int d = 0; d < 3; ++d) {
    if (d < 3 - P.ndim) {
      T.cnt[d] = 1;
      T.tile[d][0] = 0;
      continue;
    }
    const TapRange r =
        tap_range(t[d - (3 - P.ndim)], P.half[d], P.ntap[d], P.n[d], P.pad);
    if (r.count == 0) return false;
    int c = 0;
    for (int k = 0; k < r.count; ++k) {
      const int64_t g = P.pad == TNL_PAD_CIRCULAR ? mod_pos(r.g0 + k, P.n[d])
                                                  : r.g0 + k;
      const int tt = (int)(g / P.gtile[d]);
      bool seen = false;
      for (int j = 0; j < c; ++j) seen |= T.tile[d][j] == tt;
      if (!seen && c < 3) T.tile[d][c++] = tt;
    }
    T.cnt[d] = c;
  }
  return true;
}

__global__ void zero_ints_kernel(int* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ __launch_bounds__(PLAN_THREADS) void count_kernel(
    const float* __restrict__ locs, PlanDev P, int* __restrict__ counts) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < P.npts;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = base + threadIdx.x;
    const bool active = p < P.npts;
    float t[3] = {0.f, 0.f, 0.f};
    if (active)
      for (int d = 0; d < P.ndim; ++d) t[d] = locs[p * P.ndim + d];
    const int bin = active ? u_bin_of(t, P) : 0;
    agg_inc(counts, bin, active);
  }
}

__global__ __launch_bounds__(PLAN_THREADS) void scatter_kernel(
    const float* __restrict__ locs, PlanDev P, const int* __restrict__ starts,
    int* __restrict__ cursor, int* __restrict__ perm,
    float* __restrict__ slocs) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < P.npts;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = base + threadIdx.x;
    const bool active = p < P.npts;
    float t[3] = {0.f, 0.f, 0.f};
    if (active)
      for (int d = 0; d < P.ndim; ++d) t[d] = locs[p * P.ndim + d];
    const int bin = active ? u_bin_of(t, P) : 0;
    const int slot = agg_inc(cursor, bin, active);
    if (active) {
      const int pos = starts[bin] + slot;
      perm[pos] = (int)p;
      for (int d = 0; d < P.ndim; ++d) slocs[(int64_t)pos * P.ndim + d] = t[d];
    }
  }
}

// Touch counting / scattering walk the points in sorted order, so each
// tile's list comes out grouped by interpolation bin.
template <bool SCATTER>
__global__ __launch_bounds__(PLAN_THREADS) void touch_kernel(
    const float* __restrict__ slocs, PlanDev P, int* __restrict__ gctr,
    const int* __restrict__ gstarts, int* __restrict__ tlist) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < P.npts;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = base + threadIdx.x;
    const bool active = q < P.npts;
    float t[3] = {0.f, 0.f, 0.f};
    if (active)
      for (int d = 0; d < P.ndim; ++d) t[d] = slocs[q * P.ndim + d];
    Touch T;
    const bool live = active && touches_of(t, P, T);
    const int c0 = live ? T.cnt[0] : 0, c1 = live ? T.cnt[1] : 0,
              c2 = live ? T.cnt[2] : 0;
    const int ntouch = c0 * c1 * c2;
    // every lane walks the same 27 slots so ballots stay wave-uniform
    for (int s = 0; s < 27; ++s) {
      const bool act = s < ntouch;
      int tile = 0;
      if (act) {
        const int a = s / (c1 * c2), rem = s - a * (c1 * c2);
        const int b = rem / c2, c = rem - b * c2;
        tile = (T.tile[0][a] * P.gntiles[1] + T.tile[1][b]) * P.gntiles[2] +
               T.tile[2][c];
      }
      if (__ballot(act) == 0) break;
      const int slot = agg_inc(gctr, tile, act);
      if (SCATTER && act) tlist[gstarts[tile] + slot] = (int)q;
    }
  }
}

// Exclusive scan of n ints (one workgroup): out[0..n], out[n] = total.
// in may be given as a transform: k = ceil(in / div) (min 1 if min1).
__global__ __launch_bounds__(1024) void scan_kernel(const int* __restrict__ in,
                                                    int* __restrict__ out,
                                                    int64_t n, int div,
                                                    int min1) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int64_t chunk = (n + 1023) / 1024;
  const int64_t lo = tid * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  auto val = [&](int64_t i) {
    int v = in[i];
    if (div > 1) v = (v + div - 1) / div;
    if (min1 && v < 1) v = 1;
    return v;
  };
  int s = 0;
  for (int64_t i = lo; i < hi; ++i) s += val(i);
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off *= 2) {
    const int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = tid > 0 ? part[tid - 1] : 0;
  for (int64_t i = lo; i < hi; ++i) {
    out[i] = run;
    run += val(i);
  }
  if (tid == 1023) out[n] = part[1023];
}

// Interpolation subproblems: <= max_sub points of one bin each.  The dead
// bin (index u_nbins: points with no taps) gets subproblems too, so that
// interpolation writes their zero outputs.
__global__ void sub_fill_kernel(const int* __restrict__ counts,
                                const int* __restrict__ starts,
                                const int* __restrict__ suboff, PlanDev P,
                                int4* __restrict__ sub) {
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       b <= P.u_nbins; b += (int64_t)gridDim.x * blockDim.x) {
    const int c = counts[b];
    const int o = suboff[b];
    for (int j = 0; j * P.max_sub < c; ++j) {
      const int cnt = min(P.max_sub, c - j * P.max_sub);
      sub[o + j] = make_int4((int)b, starts[b] + j * P.max_sub, cnt, 0);
    }
  }
}

// Gridding work items: every tile gets >= 1 item (empty tiles write zeros);
// tiles with more than g_cap touches are split and flagged shared (w = 1).
__global__ void item_fill_kernel(const int* __restrict__ gcounts,
                                 const int* __restrict__ gstarts,
                                 const int* __restrict__ itemoff, PlanDev P,
                                 int4* __restrict__ items) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       t < P.g_ntiles_total; t += (int64_t)gridDim.x * blockDim.x) {
    const int c = gcounts[t];
    const int o = itemoff[t];
    const int k = c <= P.g_cap ? 1 : (c + P.g_cap - 1) / P.g_cap;
    for (int j = 0; j < k; ++j) {
      const int cnt = max(0, min(P.g_cap, c - j * P.g_cap));
      items[o + j] = make_int4((int)t, gstarts[t] + j * P.g_cap, cnt,
                               k > 1 ? 1 : 0);
    }
  }
}

// Gridding work items, largest first: a stable counting sort of the items by
// touch-count class (count / 512, at most 15), classes in descending order,
// the tile order kept inside a class.  The persistent gridding kernels walk
// the items in this order, so the dense tiles (the centre of a radial
// trajectory) start first instead of last; a finer, full sort by count
// scrambles the tile order of uniform trajectories for no gain.  3-D plans
// only: the batched 2-D problems ran slower with it.  One workgroup: per
// chunk of 1024 items, a ballot per class gives each item its rank among
// the chunk's items of its class.
constexpr int ITEM_CLASSES = 16;

__device__ __forceinline__ int item_class(int count) {
  const int c = count / 512;
  return c < ITEM_CLASSES - 1 ? c : ITEM_CLASSES - 1;
}

__global__ __launch_bounds__(1024) void item_order_kernel(const int4* __restrict__ items,
                                                          const int* __restrict__ nitems_ptr,
                                                          int4* __restrict__ out) {
  __shared__ int base[ITEM_CLASSES];
  __shared__ int wtot[1024 / 64][ITEM_CLASSES];
  const int n = *nitems_ptr;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < ITEM_CLASSES) base[tid] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += 1024) atomicAdd(&base[item_class(items[i].z)], 1);
  __syncthreads();
  if (tid == 0) {  // exclusive offsets, largest class first
    int acc = 0;
    for (int k = ITEM_CLASSES - 1; k >= 0; --k) {
      const int h = base[k];
      base[k] = acc;
      acc += h;
    }
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int i = c0 + tid;
    const bool valid = i < n;
    const int4 it = valid ? items[i] : make_int4(0, 0, 0, 0);
    const int k = valid ? item_class(it.z) : -1;
    int rank = 0;
    for (int b = 0; b < ITEM_CLASSES; ++b) {
      const uint64_t m = __ballot(k == b);
      if (lane == 0) wtot[wave][b] = __popcll(m);
      if (k == b) rank = __popcll(m & below);
    }
    __syncthreads();
    if (valid) {
      int pos = base[k] + rank;
      for (int w = 0; w < wave; ++w) pos += wtot[w][k];
      out[pos] = it;
    }
    __syncthreads();
    if (tid < ITEM_CLASSES) {
      int t = 0;
      for (int w = 0; w < 1024 / 64; ++w) t += wtot[w][tid];
      base[tid] += t;
    }
    __syncthreads();
  }
}

__global__ void item_copy_kernel(const int4* __restrict__ src, const int* __restrict__ nitems_ptr,
                                 int4* __restrict__ dst) {
  const int n = *nitems_ptr;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Bitonic stages below run one thread per compare pair (n/2 of them) rather
// than per element: pair p of a stage with distance j (a power of two) is
// (i, i | j), where i is p with a zero bit inserted at j's position, so
// every lane does work at every distance.
__device__ __forceinline__ int bitonic_low(int p, int j) {
  return ((p & ~(j - 1)) << 1) | (p & (j - 1));
}

// Interpolation gathers with lane = point: in one LDS instruction every lane
// reads its point's window row at the same tap offset, so two lanes whose
// rows start in the same bank (complex cells: 8-byte words, 32 bank pairs)
// serialise.  Each subproblem's points are reordered so that consecutive
// points start in different banks: by rank within their bank class (rank
// by original index), then by class.  The order depends only on the
// trajectory, so the plan stays deterministic.  One workgroup per
// subproblem; subproblems above CAP points (a max_sub tuned above 2048)
// keep the bin order.  CAP sizes the LDS buffers (24 bytes a point): plans
// with max_sub <= 512 take the 12 KB form, which keeps more workgroups
// resident per CU than the 48 KB one; both give the same order.
template <int CAP>
__global__ __launch_bounds__(256) void sub_bank_order_kernel(const int4* __restrict__ sub,
                                                             const int* __restrict__ nsub_ptr,
                                                             PlanDev P, int* __restrict__ perm,
                                                             float* __restrict__ slocs) {
  static_assert(CAP <= 2048, "the key keeps the point's slot in 11 bits");
  __shared__ unsigned long long key[CAP];
  __shared__ float s_loc[CAP * 3];
  __shared__ int s_perm[CAP];
  __shared__ int ccount[32], cstart[32];
  const int nsub = *nsub_ptr;
  int win[3];
  for (int d = 0; d < 3; ++d) win[d] = P.utile[d] + P.ntap[d] - 1;
  for (int si = blockIdx.x; si < nsub; si += gridDim.x) {
    const int4 sp = sub[si];
    const int bin = sp.x, start = sp.y, count = sp.z;
    if (bin >= P.u_nbins || count < 2 || count > CAP) continue;
    int origin[3];
    {
      int r = bin;
      for (int d = 2; d >= 0; --d) {
        origin[d] = (r % P.untiles[d]) * P.utile[d];
        r /= P.untiles[d];
      }
    }
    if (threadIdx.x < 32) ccount[threadIdx.x] = 0;
    __syncthreads();
    int n2 = 2;
    while (n2 < count) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      unsigned long long k = ~0ull;
      if (i < count) {
        const int64_t q = start + i;
        int lo[3] = {0, 0, 0};
        for (int d = 3 - P.ndim; d < 3; ++d) {
          const float t = slocs[q * P.ndim + d - (3 - P.ndim)];
          s_loc[i * 3 + d] = t;
          const TapRange r = tap_range(t, P.half[d], P.ntap[d], P.n[d], P.pad);
          const int64_t gw = P.pad == TNL_PAD_CIRCULAR ? mod_pos(r.g0, P.n[d]) : r.g0;
          lo[d] = (int)(gw - origin[d]);
        }
        const int pid = perm[q];
        s_perm[i] = pid;
        const unsigned cls = (unsigned)((lo[0] * win[1] + lo[1]) * win[2] + lo[2]) & 31u;
        atomicAdd(&ccount[cls], 1);
        k = ((unsigned long long)cls << 43) | ((unsigned long long)(unsigned)pid << 11) |
            (unsigned)i;
      }
      key[i] = k;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int acc = 0;
      for (int c = 0; c < 32; ++c) {
        cstart[c] = acc;
        acc += ccount[c];
      }
    }
    // sort by (class, original index), then re-key by (rank, class)
    for (int pass = 0; pass < 2; ++pass) {
      for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int p = threadIdx.x; p < (n2 >> 1); p += blockDim.x) {
            const int i = bitonic_low(p, j), l = i | j;
            const unsigned long long a = key[i], b = key[l];
            if ((a > b) == ((i & k) == 0)) {
              key[i] = b;
              key[l] = a;
            }
          }
          __syncthreads();
        }
      }
      if (pass == 0) {
        for (int i = threadIdx.x; i < count; i += blockDim.x) {
          const unsigned long long k = key[i];
          const int cls = (int)(k >> 43);
          const unsigned long long rank = (unsigned long long)(i - cstart[cls]);
          key[i] = (rank << 16) | ((unsigned long long)cls << 11) | (k & 2047ull);
        }
        __syncthreads();
      }
    }
    for (int j = threadIdx.x; j < count; j += blockDim.x) {
      const int i = (int)(key[j] & 2047ull);
      const int64_t q = start + j;
      perm[q] = s_perm[i];
      for (int d = 3 - P.ndim; d < 3; ++d) slocs[q * P.ndim + d - (3 - P.ndim)] = s_loc[i * 3 + d];
    }
    __syncthreads();
  }
}

// Matrix-core gridding (3-D 8x8x8 / 2-D 16x32 tiles) accumulates blocks of 16 cells
// for groups of 4 consecutive touches of a work item and skips the blocks
// no touch of the group reaches.  Ordering each item's touches by their
// block mask puts touches with equal masks side by side, so a group's
// union of blocks is close to one touch's.  Ties go by the point's index in
// the caller's order (not its sorted index, which the atomics of the point
// scatter decide), so the summation order of an item is the same for every
// plan of a trajectory.  One workgroup per item: bitonic sort of
// (mask << 32 | original index) keys in LDS, each carrying its sorted
// index; items above CAP touches (only with a g_cap tuned above 4096)
// keep the scatter order.  CAP sizes the LDS key buffer (12 bytes a key):
// plans with g_cap <= 2048 take the 24 KB form, twice as many resident
// workgroups per CU as the 48 KB one.  Keys are unique within an item, so
// both forms give the same order.
// Union of the block masks (key bits 32..63) of the group of 4 touches
// holding touch i (groups start at the item's first touch).
__device__ __forceinline__ unsigned group_mask(const unsigned long long* key, int i, int count) {
  const int i0 = i & ~3;
  unsigned m = 0;
  for (int j = i0; j < i0 + 4 && j < count; ++j) m |= (unsigned)(key[j] >> 32);
  return m;
}

template <int CAP>
__global__ __launch_bounds__(256) void item_sort_kernel(const int4* __restrict__ items,
                                                        const int* __restrict__ nitems_ptr,
                                                        const float* __restrict__ slocs,
                                                        const int* __restrict__ perm,
                                                        PlanDev P, int* __restrict__ tlist,
                                                        int2* __restrict__ tpm) {
  __shared__ unsigned long long key[CAP];
  __shared__ int val[CAP];
  const int nitems = *nitems_ptr;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int4 item = items[it];
    const int start = item.y, count = item.z;
    if (count < 1) continue;
    int origin[3];
    {
      int r = item.x;
      for (int d = 2; d >= 0; --d) {
        origin[d] = (r % P.gntiles[d]) * P.gtile[d];
        r /= P.gntiles[d];
      }
    }
    // (block mask << 32 | caller index) of a touch
    auto key_of = [&](int q) {
      unsigned mask;
      if (P.ndim == 3) {
        unsigned m[3];
        for (int d = 0; d < 3; ++d)
          m[d] = tile_cells(slocs[(int64_t)q * 3 + d], P.half[d], P.ntap[d], P.n[d], P.pad,
                            origin[d], 8);
        mask = tile_blocks32(m[0], m[1], m[2]);
      } else {
        const unsigned my = tile_cells(slocs[(int64_t)q * 2], P.half[1], P.ntap[1], P.n[1],
                                       P.pad, origin[1], 16);
        const unsigned mz = tile_cells(slocs[(int64_t)q * 2 + 1], P.half[2], P.ntap[2],
                                       P.n[2], P.pad, origin[2], 32);
        mask = tile_blocks32_2d(my, mz);
      }
      return ((unsigned long long)mask << 32) | (unsigned)perm[q];
    };
    if (count > CAP) {  // scatter order kept; touch pairs still written
      if (tpm)
        for (int g0 = 0; g0 < count; g0 += 4 * (int)blockDim.x) {
          // one thread per group of 4 touches
          const int i0 = g0 + 4 * (int)threadIdx.x;
          if (i0 < count) {
            unsigned long long k[4];
            unsigned m = 0;
            const int n = count - i0 < 4 ? count - i0 : 4;
            for (int j = 0; j < n; ++j) {
              k[j] = key_of(tlist[start + i0 + j]);
              m |= (unsigned)(k[j] >> 32);
            }
            for (int j = 0; j < n; ++j)
              tpm[start + i0 + j] = make_int2((int)(unsigned)k[j], (int)m);
          }
        }
      continue;
    }
    int n2 = 2;
    while (n2 < count) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      unsigned long long k = ~0ull;
      int q = 0;
      if (i < count) {
        q = tlist[start + i];
        k = key_of(q);
      }
      key[i] = k;
      val[i] = q;
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int p = threadIdx.x; p < (n2 >> 1); p += blockDim.x) {
          const int i = bitonic_low(p, j), l = i | j;
          const unsigned long long a = key[i], b = key[l];
          if ((a > b) == ((i & k) == 0)) {
            key[i] = b;
            key[l] = a;
            const int t = val[i];
            val[i] = val[l];
            val[l] = t;
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
      tlist[start + i] = val[i];
      if (tpm) tpm[start + i] = make_int2((int)(unsigned)key[i], (int)group_mask(key, i, count));
    }
    __syncthreads();
  }
}

PlanDev make_plan_dev(const tnl_plan_layout& L) {
  PlanDev P;
  std::memset(&P, 0, sizeof(P));
  P.ndim = L.ndim;
  for (int d = 0; d < 3; ++d) {
    P.n[d] = 1;
    P.half[d] = 0.5f;
    P.ntap[d] = 1;
    P.utile[d] = L.u_tile[d];
    P.untiles[d] = L.u_ntiles[d];
    P.gtile[d] = L.g_tile[d];
    P.gntiles[d] = L.g_ntiles[d];
  }
  for (int gd = 3 - L.ndim; gd < 3; ++gd) {
    P.n[gd] = L.grid_size[gd];
    P.half[gd] = L.width[gd] * 0.5f;
    P.ntap[gd] = L.ntaps[gd];
  }
  P.pad = L.pad_mode;
  P.u_nbins = L.u_nbins;
  P.g_ntiles_total = L.g_ntiles_total;
  P.npts = L.npts;
  P.max_sub = (int)L.max_sub;
  P.g_cap = (int)L.g_cap;
  return P;
}

static unsigned grid_for(int64_t n, int64_t cap = 4096) {
  int64_t b = (n + PLAN_THREADS - 1) / PLAN_THREADS;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_plan_layout_query(int64_t npts, int ndim,
                                     const int64_t* grid_size,
                                     const float* width, int pad_mode,
                                     const tnl_tuning* tuning,
                                     tnl_plan_layout* out) {
  if (!out || !grid_size || !width) return fail(TNL_EINVAL, "null pointer");
  return plan_layout(npts, ndim, grid_size, width, pad_mode, tuning, out);
}

extern "C" int tnl_plan_build(const float* locs, const tnl_plan_layout* layout,
                              void* workspace, int64_t workspace_bytes,
                              void* stream) {
  if (!layout) return fail(TNL_EINVAL, "null layout");
  const tnl_plan_layout L = *layout;
  if (L.ndim < 1 || L.ndim > 3 || L.npts < 0 || L.bytes <= 0)
    return fail(TNL_EINVAL, "invalid plan layout");
  const int64_t npts = L.npts;
  if (!workspace) return fail(TNL_EINVAL, "null workspace");
  if (workspace_bytes < L.bytes)
    return fail(TNL_EWORKSPACE, "plan workspace too small");
  if (npts > 0 && !locs) return fail(TNL_EINVAL, "null locs");
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  int* counts = (int*)(ws + L.off_counts);
  int* starts = (int*)(ws + L.off_starts);
  int* cursor = (int*)(ws + L.off_cursor);
  int* perm = (int*)(ws + L.off_perm);
  float* slocs = (float*)(ws + L.off_slocs);
  int* suboff = (int*)(ws + L.off_suboff);
  int4* sub = (int4*)(ws + L.off_sub);
  int* gcounts = (int*)(ws + L.off_gcounts);
  int* gstarts = (int*)(ws + L.off_gstarts);
  int* gcursor = (int*)(ws + L.off_gcursor);
  int* tlist = (int*)(ws + L.off_tlist);
  int* itemoff = (int*)(ws + L.off_itemoff);
  int4* items = (int4*)(ws + L.off_items);
  int2* tpm = L.off_tpm ? (int2*)(ws + L.off_tpm) : nullptr;
  const PlanDev P = make_plan_dev(L);
  const int64_t nb1 = L.u_nbins + 1;

  zero_ints_kernel<<<grid_for(nb1), PLAN_THREADS, 0, s>>>(counts, nb1);
  zero_ints_kernel<<<grid_for(nb1), PLAN_THREADS, 0, s>>>(cursor, nb1);
  zero_ints_kernel<<<grid_for(L.g_ntiles_total), PLAN_THREADS, 0, s>>>(
      gcounts, L.g_ntiles_total);
  zero_ints_kernel<<<grid_for(L.g_ntiles_total), PLAN_THREADS, 0, s>>>(
      gcursor, L.g_ntiles_total);
  TNL_LAUNCH_CHECK("plan zero");
  if (npts > 0) {
    count_kernel<<<grid_for(npts), PLAN_THREADS, 0, s>>>(locs, P, counts);
    TNL_LAUNCH_CHECK("plan count");
  }
  scan_kernel<<<1, 1024, 0, s>>>(counts, starts, nb1, 1, 0);
  TNL_LAUNCH_CHECK("plan scan");
  if (npts > 0) {
    scatter_kernel<<<grid_for(npts), PLAN_THREADS, 0, s>>>(
        locs, P, starts, cursor, perm, slocs);
    TNL_LAUNCH_CHECK("plan scatter");
    touch_kernel<false><<<grid_for(npts), PLAN_THREADS, 0, s>>>(
        slocs, P, gcounts, nullptr, nullptr);
    TNL_LAUNCH_CHECK("plan touch count");
  }
  // subproblems: k_b = ceil(count_b / max_sub), dead bin included
  scan_kernel<<<1, 1024, 0, s>>>(counts, suboff, nb1, (int)L.max_sub, 0);
  sub_fill_kernel<<<grid_for(nb1), PLAN_THREADS, 0, s>>>(
      counts, starts, suboff, P, sub);
  TNL_LAUNCH_CHECK("plan subproblems");
  if (npts > 0) {
    const int64_t nwg = L.sub_capacity < 8192 ? L.sub_capacity : 8192;
    if (L.max_sub <= 512)
      sub_bank_order_kernel<512><<<(unsigned)nwg, 256, 0, s>>>(sub, suboff + L.u_nbins + 1, P,
                                                              perm, slocs);
    else
      sub_bank_order_kernel<2048><<<(unsigned)nwg, 256, 0, s>>>(sub, suboff + L.u_nbins + 1, P,
                                                               perm, slocs);
    TNL_LAUNCH_CHECK("plan subproblem order");
  }
  scan_kernel<<<1, 1024, 0, s>>>(gcounts, gstarts, L.g_ntiles_total, 1, 0);
  if (npts > 0) {
    touch_kernel<true><<<grid_for(npts), PLAN_THREADS, 0, s>>>(
        slocs, P, gcursor, gstarts, tlist);
    TNL_LAUNCH_CHECK("plan touch scatter");
  }
  scan_kernel<<<1, 1024, 0, s>>>(gcounts, itemoff, L.g_ntiles_total,
                                 (int)L.g_cap, 1);
  item_fill_kernel<<<grid_for(L.g_ntiles_total), PLAN_THREADS, 0, s>>>(
      gcounts, gstarts, itemoff, P, items);
  TNL_LAUNCH_CHECK("plan items");
  if (npts > 0 && L.grid_accum == TNL_GRID_ACCUM_MFMA) {
    const int64_t nwg = L.item_capacity < 8192 ? L.item_capacity : 8192;
    if (L.g_cap <= 2048)
      item_sort_kernel<2048><<<(unsigned)nwg, 256, 0, s>>>(items, itemoff + L.g_ntiles_total,
                                                          slocs, perm, P, tlist, tpm);
    else
      item_sort_kernel<4096><<<(unsigned)nwg, 256, 0, s>>>(items, itemoff + L.g_ntiles_total,
                                                          slocs, perm, P, tlist, tpm);
    TNL_LAUNCH_CHECK("plan item sort");
  }
  if (npts > 0 && L.ndim == 3) {
    int4* tmp = (int4*)(ws + L.off_items_tmp);
    item_order_kernel<<<1, 1024, 0, s>>>(items, itemoff + L.g_ntiles_total, tmp);
    const int64_t nwg = (L.item_capacity + 255) / 256;
    item_copy_kernel<<<(unsigned)(nwg < 4096 ? (nwg < 1 ? 1 : nwg) : 4096), 256, 0, s>>>(
        tmp, itemoff + L.g_ntiles_total, items);
    TNL_LAUNCH_CHECK("plan item order");
  }
  return TNL_OK;
}
