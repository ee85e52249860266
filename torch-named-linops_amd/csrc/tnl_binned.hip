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

template <int NDIM, bool CPLX, int CC, int NT>
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
  if (bin == B.nbins) {  // points without taps: zero output
    T* __restrict__ o = reinterpret_cast<T*>(out);
    for (int i = threadIdx.x; i < count; i += BIN_THREADS) {
      const int p = perm[start + i];
      for (int c = 0; c < nc; ++c) o[(c0 + c) * B.npts + p] = C::zero();
    }
    return;
  }

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
  // ---- stage window: channel planes [c][cell] in LDS (lanes gathering
  // different cells then fall in different banks), coalesced along z in
  // HBM.  A table of row bases (one per (channel, x, y) row of the window)
  // keeps the per-element index math to one division by the row length.
  const T* __restrict__ v = reinterpret_cast<const T*>(vals);
  const int w2 = B.win[2];
  const int nrows = CC * B.win[0] * B.win[1];
  int64_t* rowbase = reinterpret_cast<int64_t*>(
      lds_raw + ((size_t)CC * B.wcells * sizeof(T) + 7) / 8 * 8);
  for (int r = threadIdx.x; r < nrows; r += BIN_THREADS) {
    const int c = r / (B.win[0] * B.win[1]);
    const int rem = r - c * (B.win[0] * B.win[1]);
    const int wx = rem / B.win[1], wy = rem - wx * B.win[1];
    int64_t gx = origin[0] + wx, gy = origin[1] + wy;
    bool ok = c < nc;
    if (G.pad == TNL_PAD_CIRCULAR) {
      gx = mod_pos(gx, G.n[0]);
      gy = mod_pos(gy, G.n[1]);
    } else {
      ok = ok && gx < G.n[0] && gy < G.n[1];
    }
    rowbase[r] = ok ? (c0 + c) * G.ncells + (gx * G.n[1] + gy) * G.n[2] : -1;
  }
  __syncthreads();
  const float inv_w2 = 1.0f / (float)w2;
  const int total = nrows * w2;
  for (int idx = threadIdx.x; idx < total; idx += BIN_THREADS) {
    int r = (int)(((float)idx + 0.5f) * inv_w2);
    if (r * w2 > idx) --r;              // correct the float quotient
    else if ((r + 1) * w2 <= idx) ++r;
    const int wz = idx - r * w2;
    int64_t gz = origin[2] + wz;
    const int64_t base = rowbase[r];
    bool ok = base >= 0;
    if (G.pad == TNL_PAD_CIRCULAR) {
      if (gz >= G.n[2]) gz = mod_pos(gz, G.n[2]);
    } else {
      ok = ok && gz < G.n[2];
    }
    T val = C::zero();
    if (ok) val = v[base + gz];
    // channel c of row r lives in plane c; rows of one channel are
    // contiguous, so the LDS index is simply idx
    lds[idx] = val;
  }
  __syncthreads();

  // ---- gather: one point per lane.  The x taps are a runtime loop (their
  // window values computed on the fly), y and z taps are unrolled to NT:
  // the loop body stays small enough for the instruction cache.
  for (int i = threadIdx.x; i < count; i += BIN_THREADS) {
    const int q = start + i;
    const int p = perm[q];
    float t[3] = {0.f, 0.f, 0.f};
    TapRange r[3];
    int lo[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      r[d].g0 = 0;
      r[d].count = 1;
      lo[d] = 0;
      if (d >= 3 - NDIM) {
        t[d] = slocs[(int64_t)q * NDIM + d - (3 - NDIM)];
        r[d] = tap_range(t[d], G.half[d], G.ntap[d], G.n[d], G.pad);
        const int64_t gw =
            G.pad == TNL_PAD_CIRCULAR ? mod_pos(r[d].g0, G.n[d]) : r[d].g0;
        lo[d] = (int)(gw - origin[d]);
      }
    }
    float wy[NT], wz[NT];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      wy[k] = 0.0f;
      wz[k] = 0.0f;
      if (NDIM >= 2 && k < r[1].count)
        wy[k] = window(G.kernel,
                       fminf(fabsf((float)(r[1].g0 + k) - t[1]) * G.inv_half[1], 1.0f),
                       G.beta);
      if (k < r[2].count)
        wz[k] = window(G.kernel,
                       fminf(fabsf((float)(r[2].g0 + k) - t[2]) * G.inv_half[2], 1.0f),
                       G.beta);
    }
    if (NDIM < 2) wy[0] = 1.0f;
    T acc[CC];
#pragma unroll
    for (int c = 0; c < CC; ++c) acc[c] = C::zero();
#pragma unroll 1
    for (int kx = 0; kx < r[0].count; ++kx) {
      float wx = 1.0f;
      if (NDIM >= 3)
        wx = window(G.kernel,
                    fminf(fabsf((float)(r[0].g0 + kx) - t[0]) * G.inv_half[0], 1.0f),
                    G.beta);
      const int xrow = (lo[0] + kx) * B.win[1];
#pragma unroll
      for (int ky = 0; ky < NT; ++ky) {
        if (ky >= r[1].count) break;
        const float wxy = wx * wy[ky];
        const T* rowp = lds + (xrow + lo[1] + ky) * B.win[2] + lo[2];
#pragma unroll
        for (int kz = 0; kz < NT; ++kz) {
          if (kz >= r[2].count) break;
          const float wt = wxy * wz[kz];
#pragma unroll
          for (int c = 0; c < CC; ++c) C::fma(acc[c], wt, rowp[c * B.wcells + kz]);
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

// Gridding, "lanes = taps": one wave per workgroup owns a tile accumulator
// in LDS (channel planes, z rows padded by one cell against bank
// conflicts).  Touches are processed in batches of 64: first lanes = points
// compute each point's first tap, tap counts and 1-D window values into LDS;
// then the wave walks the batch point by point with lanes = taps (a 4x4x4
// block, plus a second pass for the fifth tap layer of points on integer
// coordinates).  Within one instruction every lane addresses a different
// cell and only this wave touches the tile, so the accumulation is a plain
// read-modify-write: no LDS atomics (which run about one lane per clock).
constexpr int GW = 64;  // threads per gridding workgroup

template <int NDIM, bool CPLX, int CC, int NT>
__global__ __launch_bounds__(GW) void grid_binned_kernel(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    InterpGeom G, BinGeom B, const int* __restrict__ perm,
    const float* __restrict__ slocs, const int* __restrict__ tlist,
    const int4* __restrict__ items, const int* __restrict__ nitems_ptr) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  constexpr int COMP = CPLX ? 2 : 1;
  constexpr int NP = CC * COMP;  // accumulator planes
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int zp = B.tile[2] + 1;                   // padded z row
  const int pcells = B.tile[0] * B.tile[1] * zp;  // cells per plane
  float* acc = reinterpret_cast<float*>(lds_raw);  // [NP][pcells]
  // point table of the current batch (lanes = points fill it)
  int* s_lo = reinterpret_cast<int*>(acc + NP * pcells);  // [3][GW] first tap - origin
  int* s_cnt = s_lo + 3 * GW;                              // [3][GW]
  float* s_w = reinterpret_cast<float*>(s_cnt + 3 * GW);   // [3][NT][GW]
  float* s_v = s_w + 3 * NT * GW;                          // [NP][GW]
  const int lane = threadIdx.x;
  const int nitems = *nitems_ptr;
  const T* __restrict__ v = reinterpret_cast<const T*>(vals);

  // This lane's taps.  3-D: main pass (k0,k1,k2) in [0,4)^3 and a second
  // pass over the 61 taps of [0,5)^3 with some index 4 (points on integer
  // coordinates).  2-D: one pass over (k1,k2) in [0,8)^2.
  int mk[3], sk[3];
  bool shell_ok = false;
  if (NDIM >= 3) {
    mk[0] = lane >> 4; mk[1] = (lane >> 2) & 3; mk[2] = lane & 3;
    sk[0] = sk[1] = sk[2] = 0;
    int seen = 0;
    for (int a = 0; a < 5; ++a)
      for (int b2 = 0; b2 < 5; ++b2)
        for (int c2 = 0; c2 < 5; ++c2)
          if (a == 4 || b2 == 4 || c2 == 4) {
            if (seen == lane) { sk[0] = a; sk[1] = b2; sk[2] = c2; shell_ok = true; }
            ++seen;
          }
  } else {
    mk[0] = 0; mk[1] = lane >> 3; mk[2] = lane & 7;
    sk[0] = sk[1] = sk[2] = 0;
  }

  for (int64_t it = xcd_remap(blockIdx.x, gridDim.x); it < nitems;
       it += gridDim.x) {
    const int4 item = items[it];
    const int tile = item.x, start = item.y, count = item.z;
    const bool shared = item.w != 0;
    int origin[3];
    {
      int r = tile;
      for (int d = 2; d >= 0; --d) {
        origin[d] = (r % B.ntiles[d]) * B.tile[d];
        r /= B.ntiles[d];
      }
    }
    for (int64_t c0 = 0; c0 < nbatch; c0 += CC) {
      const int nc = (int)(nbatch - c0 < CC ? nbatch - c0 : CC);
      for (int i = lane; i < NP * pcells; i += GW) acc[i] = 0.0f;
      for (int b0 = 0; b0 < count; b0 += GW) {
        const int nb = count - b0 < GW ? count - b0 : GW;
        __syncthreads();
        // ---- lanes = points: first tap, counts, windows and values
        if (lane < nb) {
          const int q = tlist[start + b0 + lane];
          const int p = perm[q];
#pragma unroll
          for (int c = 0; c < CC; ++c) {
            const T x = c < nc ? v[(c0 + c) * B.npts + p] : C::zero();
            if constexpr (CPLX) {
              s_v[(2 * c) * GW + lane] = x.x;
              s_v[(2 * c + 1) * GW + lane] = x.y;
            } else {
              s_v[c * GW + lane] = x;
            }
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            int lo = 0, cnt = 1;
            float t = 0.f;
            TapRange r;
            r.g0 = 0;
            r.count = 1;
            if (d >= 3 - NDIM) {
              t = slocs[(int64_t)q * NDIM + d - (3 - NDIM)];
              r = tap_range(t, G.half[d], G.ntap[d], G.n[d], G.pad);
              const int64_t gw =
                  G.pad == TNL_PAD_CIRCULAR ? mod_pos(r.g0, G.n[d]) : r.g0;
              lo = (int)(gw - origin[d]);
              cnt = r.count;
            }
            s_lo[d * GW + lane] = lo;
            s_cnt[d * GW + lane] = cnt;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              float w = (d < 3 - NDIM && k == 0) ? 1.0f : 0.0f;
              if (d >= 3 - NDIM && k < r.count)
                w = window(G.kernel,
                           fminf(fabsf((float)(r.g0 + k) - t) * G.inv_half[d], 1.0f),
                           G.beta);
              s_w[(d * NT + k) * GW + lane] = w;
            }
          }
        }
        __syncthreads();
        // ---- lanes = taps: walk the batch point by point.  The next
        // point's table entries are loaded before this point's
        // read-modify-write so their LDS latency overlaps it.
        int nx_lo[3], nx_n[3];
        float nx_w[3];
        float nx_v[NP];
        auto fetch = [&](int j, const int* k) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            nx_lo[d] = s_lo[d * GW + j];
            nx_n[d] = s_cnt[d * GW + j];
            nx_w[d] = s_w[(d * NT + (k[d] < NT ? k[d] : 0)) * GW + j];
          }
#pragma unroll
          for (int c = 0; c < NP; ++c) nx_v[c] = s_v[c * GW + j];
        };
        fetch(0, mk);
        for (int j = 0; j < nb; ++j) {
          int lo[3], n[3];
          float w3[3], vv[NP];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            lo[d] = nx_lo[d];
            n[d] = nx_n[d];
            w3[d] = nx_w[d];
          }
#pragma unroll
          for (int c = 0; c < NP; ++c) vv[c] = nx_v[c];
          if (j + 1 < nb) fetch(j + 1, mk);
          const bool fifth = NDIM >= 3 && (n[0] > 4 || n[1] > 4 || n[2] > 4);
#pragma unroll
          for (int pass = 0; pass < (NDIM >= 3 ? 2 : 1); ++pass) {
            if (pass == 1 && !fifth) break;
            const int* k = pass ? sk : mk;
            bool ok = (pass == 0 || shell_ok) && k[0] < n[0] && k[1] < n[1] && k[2] < n[2];
            int l[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              l[d] = lo[d] + k[d];
              // footprint wrapping past the end of the grid
              if (G.pad == TNL_PAD_CIRCULAR && origin[d] + l[d] >= G.n[d])
                l[d] -= (int)G.n[d];
              ok = ok && l[d] >= 0 && l[d] < B.tile[d];
            }
            if (ok) {
              float w;
              if (pass == 0) {
                w = w3[0] * w3[1] * w3[2];
              } else {
                w = s_w[(0 * NT + k[0]) * GW + j] * s_w[(1 * NT + k[1]) * GW + j] *
                    s_w[(2 * NT + k[2]) * GW + j];
              }
              float* cp = acc + (l[0] * B.tile[1] + l[1]) * zp + l[2];
              float old[NP];
#pragma unroll
              for (int c = 0; c < NP; ++c) old[c] = cp[c * pcells];
#pragma unroll
              for (int c = 0; c < NP; ++c) cp[c * pcells] = fmaf(w, vv[c], old[c]);
            }
          }
        }
      }
      __syncthreads();
      // ---- write the tile: coalesced along z, one channel plane at a time
      T* __restrict__ o = reinterpret_cast<T*>(out);
      const int t12 = B.tile[1] * B.tile[2];
      const int total = B.tcells * nc;
      for (int idx = lane; idx < total; idx += GW) {
        const int c = idx / B.tcells;
        const int cell = idx - c * B.tcells;
        const int lx = cell / t12, rem = cell - lx * t12;
        const int ly = rem / B.tile[2], lz = rem - ly * B.tile[2];
        const int64_t gx = origin[0] + lx, gy = origin[1] + ly,
                      gz = origin[2] + lz;
        if (gx >= G.n[0] || gy >= G.n[1] || gz >= G.n[2]) continue;
        const int64_t gi = (c0 + c) * G.ncells + (gx * G.n[1] + gy) * G.n[2] + gz;
        const int pc = (lx * B.tile[1] + ly) * zp + lz;
        const float a0 = acc[(c * COMP) * pcells + pc];
        const float a1 = CPLX ? acc[(c * COMP + 1) * pcells + pc] : 0.0f;
        if (!shared) {
          if constexpr (CPLX)
            o[gi] = make_float2(a0, a1);
          else
            o[gi] = a0;
        } else {
          if constexpr (CPLX) {
            unsafeAtomicAdd(out + 2 * gi, a0);
            unsafeAtomicAdd(out + 2 * gi + 1, a1);
          } else {
            unsafeAtomicAdd(out + gi, a0);
          }
        }
      }
      __syncthreads();
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

static size_t ungrid_lds_bytes(const BinGeom& B, int cc, int comp_bytes) {
  return ((size_t)B.wcells * cc * comp_bytes + 7) / 8 * 8 +
         (size_t)cc * B.win[0] * B.win[1] * sizeof(int64_t);
}

static size_t grid_lds_bytes(const BinGeom& B, int planes, int nt) {
  const size_t pcells = (size_t)B.tile[0] * B.tile[1] * (B.tile[2] + 1);
  return planes * pcells * 4 + (6 * GW + 3 * nt * GW + planes * GW) * 4;
}

bool binned_supported(const InterpGeom& G, const tnl_plan_layout& L,
                      bool adjoint) {
  if (G.ndim < 2 || G.norm != 1) return false;
  for (int d = 0; d < 3; ++d)
    if (G.ntap[d] > 8 || G.ntap[d] != L.ntaps[d]) return false;
  if (adjoint) {
    for (int d = 3 - G.ndim; d < 3; ++d) {
      if (G.ntap[d] > 5) return false;       // gridding lanes cover 5 taps
      if (G.n[d] < G.ntap[d]) return false;  // a footprint would self-wrap
    }
    return (int64_t)grid_lds_bytes(make_bin_geom(L, true), 2, 5) <= L.lds_budget;
  }
  // at least one channel of the window must fit in LDS
  return (int64_t)ungrid_lds_bytes(make_bin_geom(L, false), 1, 8) <= L.lds_budget;
}

// Dynamic LDS above the default 64 KiB needs an explicit opt-in per kernel.
static int allow_lds(const void* fn, size_t bytes) {
  if (bytes <= 65536) return TNL_OK;
  return check_hip(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)bytes),
                   "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

template <int NDIM, bool CPLX, int CC, int NT>
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
    const size_t lds = ungrid_lds_bytes(B, CC, comp_bytes);
    if (int rc = allow_lds((const void*)ungrid_binned_kernel<NDIM, CPLX, CC, NT>, lds))
      return rc;
    ungrid_binned_kernel<NDIM, CPLX, CC, NT>
        <<<dim3((unsigned)nwg, nchunk), BIN_THREADS, lds, s>>>(
            vals, out, nbatch, G, B, perm, slocs,
            (const int4*)(ws + L.off_sub),
            (const int*)(ws + L.off_suboff) + L.u_nbins + 1);
    TNL_LAUNCH_CHECK("ungrid_binned");
  } else {
    const int* nitems = (const int*)(ws + L.off_itemoff) + L.g_ntiles_total;
    const int4* items = (const int4*)(ws + L.off_items);
    grid_zero_shared_kernel<CPLX><<<1024, 256, 0, s>>>(
        out, nbatch, G, B, items, nitems, (const int*)(ws + L.off_gstarts));
    TNL_LAUNCH_CHECK("grid_zero_shared");
    const size_t lds = grid_lds_bytes(B, CC * (CPLX ? 2 : 1), NT);
    if (int rc = allow_lds((const void*)grid_binned_kernel<NDIM, CPLX, CC, NT>, lds))
      return rc;
    // persistent: a few workgroups per CU walk the work items
    int64_t nwg = (L.item_capacity + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const int64_t cap = 256 * 16;
    if (nwg > cap) nwg = cap;
    grid_binned_kernel<NDIM, CPLX, CC, NT><<<(unsigned)nwg, GW, lds, s>>>(
        vals, out, nbatch, G, B, perm, slocs, (const int*)(ws + L.off_tlist),
        items, nitems);
    TNL_LAUNCH_CHECK("grid_binned");
  }
  return TNL_OK;
}

template <int NDIM, bool CPLX, int NT>
static int dispatch_cc(bool adjoint, const float* vals, float* out,
                       int64_t nbatch, const InterpGeom& G,
                       const tnl_plan_layout& L, const char* ws,
                       hipStream_t s) {
  const BinGeom B = make_bin_geom(L, adjoint);
  int cc;
  if (adjoint) {
    cc = 1;
    while (cc < 8 && cc < nbatch &&
           (int64_t)grid_lds_bytes(B, 2 * cc * (CPLX ? 2 : 1), NT) <= L.lds_budget)
      cc *= 2;
  } else {
    cc = 1;
    while (cc < 8 && cc < nbatch &&
           (int64_t)ungrid_lds_bytes(B, 2 * cc, CPLX ? 8 : 4) <= L.lds_budget)
      cc *= 2;
  }
  switch (cc) {
    case 8: return launch_binned<NDIM, CPLX, 8, NT>(adjoint, vals, out, nbatch, G, L, ws, s);
    case 4: return launch_binned<NDIM, CPLX, 4, NT>(adjoint, vals, out, nbatch, G, L, ws, s);
    case 2: return launch_binned<NDIM, CPLX, 2, NT>(adjoint, vals, out, nbatch, G, L, ws, s);
    default: return launch_binned<NDIM, CPLX, 1, NT>(adjoint, vals, out, nbatch, G, L, ws, s);
  }
}

template <int NDIM, bool CPLX>
static int dispatch_nt(bool adjoint, const float* vals, float* out,
                       int64_t nbatch, const InterpGeom& G,
                       const tnl_plan_layout& L, const char* ws,
                       hipStream_t s) {
  const int nt = G.ntap[1] > G.ntap[2] ? G.ntap[1] : G.ntap[2];
  if (nt <= 5) return dispatch_cc<NDIM, CPLX, 5>(adjoint, vals, out, nbatch, G, L, ws, s);
  return dispatch_cc<NDIM, CPLX, 8>(adjoint, vals, out, nbatch, G, L, ws, s);
}

int interp_binned(bool adjoint, const float* vals, float* out, int64_t nbatch,
                  const InterpGeom& G, int is_complex, const tnl_plan_layout& L,
                  const void* plan, hipStream_t s) {
  const char* ws = (const char*)plan;
  if (G.ndim == 3)
    return is_complex ? dispatch_nt<3, true>(adjoint, vals, out, nbatch, G, L, ws, s)
                      : dispatch_nt<3, false>(adjoint, vals, out, nbatch, G, L, ws, s);
  return is_complex ? dispatch_nt<2, true>(adjoint, vals, out, nbatch, G, L, ws, s)
                    : dispatch_nt<2, false>(adjoint, vals, out, nbatch, G, L, ws, s);
}

}  // namespace tnl
