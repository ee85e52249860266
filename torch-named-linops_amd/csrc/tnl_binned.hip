// Plan-driven interpolation (ungrid) and gridding kernels, 2-D and 3-D,
// norm 1 (separable weights).
//
// ungrid: one workgroup per subproblem.  A subproblem is <= max_sub points
//   of one bin.  For each chunk of CC channels the workgroup stages the
//   bin's grid tile plus a halo of ntap-1 cells per axis (wrapped, or zero
//   outside the grid under zero padding) in LDS, then each lane gathers the
//   taps of one point from LDS for all CC channels at once.  The per-point
//   weights are computed once, on the first chunk, and kept in registers
//   for the later chunks.
// grid: "owner computes".  One workgroup per (work item, channel chunk); a
//   work item is a gridding tile with the list of points touching it.  Taps
//   inside the tile are accumulated on the matrix cores (grid_mfma_kernel;
//   3-D plans with touch pairs hand over to the one-wave kernel of
//   tnl_gridwave.hip when the caller passes a scratch) or by plain LDS
//   read-modify-writes in which no two lanes or waves share a word
//   (grid_binned_kernel), then the tile is written to HBM once with
//   coalesced stores.  No zero-fill pass and no global atomics, except for
//   tiles with more than g_cap touches, which are split across workgroups
//   and combined with global float atomics into a region zeroed by a small
//   pre-pass.
// Workgroup ids are remapped so each XCD walks a contiguous range of tiles
// (neighbouring halos then share that XCD's L2).
#include <cstdlib>

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
  int gys;  // gridding accumulator: cells between y rows
  int gxs;  // gridding accumulator: cells between x planes
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

// One subproblem (<= max_sub points of one bin) of one problem, every
// channel, CC channels at a time.  Called by all threads of the workgroup.
// Each thread keeps the first tap offset, tap counts and window values of
// its first UPTS points in registers across the channel chunks, so the
// windows are evaluated once per point.
constexpr int UPTS = 2;  // cached points per thread (max_sub <= 512)

// ILV: the 3-D complex form for two channels per chunk on 8^3 bins with 5
// taps per axis (a 12^3 window): the window is staged cell-major with the
// two channels of a cell side by side (16 bytes), so each tap of a point
// is one 16-byte LDS read for both channels instead of two 8-byte reads of
// channel planes (measured on the 3-D radial workload, 8 coils, 256^3,
// 20000 x 512 points: the gather without its staging ran 1.74 -> 1.31 ms).
constexpr int ILV_W = 12;
#ifndef ILV_RY
#define ILV_RY 2
#endif

// pm: write point-major, out[p * nbatch + c] (all channels of a point side
// by side) instead of channel-major out[c * npts + p].
template <int NDIM, bool CPLX, int CC, int NT, bool ILV = false>
__device__ __forceinline__ void ungrid_sub(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    const InterpGeom& G, const BinGeom& B, const int* __restrict__ perm,
    const float* __restrict__ slocs, int4 sp, bool pm = false) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);
  const int bin = sp.x, start = sp.y, count = sp.z;
  T* __restrict__ o = reinterpret_cast<T*>(out);
  if (bin == B.nbins) {  // points without taps: zero output
    for (int i = threadIdx.x; i < count; i += BIN_THREADS) {
      const int p = perm[start + i];
      for (int64_t c = 0; c < nbatch; ++c) o[pm ? p * nbatch + c : c * B.npts + p] = C::zero();
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
  const T* __restrict__ v = reinterpret_cast<const T*>(vals);
  const int w2 = B.win[2];
  const int nrows = CC * B.win[0] * B.win[1];

  // per-point data cached across channel chunks (the weights of the
  // grid's axes only: a 2-D point's leading axis is the constant 1, 0, ...)
  constexpr int WOFF = (3 - NDIM) * NT;
  float cw[UPTS][NDIM * NT];
  int crow[UPTS], ccnt[UPTS], cpnt[UPTS];

  static_assert(!ILV || (NDIM == 3 && CPLX && CC == 2 && NT == 5), "interleaved form");
  for (int64_t c0 = 0; c0 < nbatch; c0 += CC) {
    const int nc = (int)(nbatch - c0 < CC ? nbatch - c0 : CC);
    if (c0 > 0) __syncthreads();  // previous chunk's gathers are done
    if constexpr (ILV) {
      // ---- stage window, interleaved: window cell e = (wx * 12 + wy) * 12
      // + wz holds channels c0 and c0 + 1 as one float4.  One pair of
      // z-neighbour cells per thread and pass: a 16-byte load per channel
      // (the loads of consecutive threads run along z) and two 16-byte
      // stores; a thread issues all its loads before its stores.  The
      // caller guarantees n >= 12 per axis and an even n[2]: one
      // subtraction wraps a window cell under circular padding and both
      // cells of a pair wrap together; under zero padding cells past the
      // grid are zero (tile origins are never negative).
      constexpr int NPAIR = ILV_W * ILV_W * ILV_W / 2;
      constexpr int NPASS = (NPAIR + BIN_THREADS - 1) / BIN_THREADS;
      const int n0 = (int)G.n[0], n1 = (int)G.n[1], n2 = (int)G.n[2];
      const bool circ = G.pad == TNL_PAD_CIRCULAR;
      const T* __restrict__ v0 = v + c0 * G.ncells;
      const T* __restrict__ v1 = v + (nc > 1 ? c0 + 1 : c0) * G.ncells;
      // recomputed per chunk: hoisted out of the chunk loop, the cell
      // coordinates of all passes would stay live through the gather
      int tid;
      asm volatile("v_mov_b32 %0, %1" : "=v"(tid) : "v"((int)threadIdx.x));
      float4 a[NPASS], b[NPASS];
#pragma unroll
      for (int i = 0; i < NPASS; ++i) {
        const int e = tid + i * BIN_THREADS;
        a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        b[i] = a[i];
        if (e < NPAIR) {
          const int row = e / (ILV_W / 2), wz = 2 * (e - row * (ILV_W / 2));
          const int wx = row / ILV_W, wy = row - wx * ILV_W;
          int gx = origin[0] + wx, gy = origin[1] + wy, gz = origin[2] + wz;
          bool in = true;
          if (circ) {
            gx = gx >= n0 ? gx - n0 : gx;
            gy = gy >= n1 ? gy - n1 : gy;
            gz = gz >= n2 ? gz - n2 : gz;
          } else {
            in = gx < n0 && gy < n1 && gz < n2;
          }
          if (in) {
            const int64_t cell = ((int64_t)gx * n1 + gy) * n2 + gz;
            a[i] = *reinterpret_cast<const float4*>(v0 + cell);
            if (nc > 1) b[i] = *reinterpret_cast<const float4*>(v1 + cell);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NPASS; ++i) {
        const int e = tid + i * BIN_THREADS;
        if (e < NPAIR) {
          float4* d = reinterpret_cast<float4*>(lds) + 2 * e;
          d[0] = make_float4(a[i].x, a[i].y, b[i].x, b[i].y);
          d[1] = make_float4(a[i].z, a[i].w, b[i].z, b[i].w);
        }
      }
    } else {
    // ---- stage window: channel planes [c][cell] in LDS (lanes gathering
    // different cells then fall in different banks).  One thread per
    // (channel, x, y) row of the window: the row's w2 cells are contiguous
    // in HBM and in LDS (row r of the window starts at lds[r * w2]), so a
    // row that neither wraps nor leaves the grid is copied with 16-byte
    // loads and stores; the others go cell by cell.
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
      T* dst = lds + (int64_t)r * w2;
      if (!ok) {
        for (int k = 0; k < w2; ++k) dst[k] = C::zero();
        continue;
      }
      const int64_t base = (c0 + c) * G.ncells + (gx * G.n[1] + gy) * G.n[2];
      const int64_t z0 = origin[2];
      constexpr int PER4 = 16 / (int)sizeof(T);  // cells per 16 bytes
      if (z0 + w2 <= G.n[2] && w2 % PER4 == 0 && (base + z0) % PER4 == 0) {
        const float4* src4 = reinterpret_cast<const float4*>(v + base + z0);
        float4* dst4 = reinterpret_cast<float4*>(dst);
        for (int k = 0; k < w2 / PER4; ++k) dst4[k] = src4[k];
      } else {
        for (int k = 0; k < w2; ++k) {
          int64_t gz = z0 + k;
          bool in = true;
          if (G.pad == TNL_PAD_CIRCULAR) {
            if (gz >= G.n[2]) gz = mod_pos(gz, G.n[2]);
          } else {
            in = gz < G.n[2];
          }
          dst[k] = in ? v[base + gz] : C::zero();
        }
      }
    }
    }
    __syncthreads();

    // ---- gather: one point per lane
    int slot = 0;
    for (int i = threadIdx.x; i < count; i += BIN_THREADS, ++slot) {
      float w[3 * NT];
      int rowoff, cnt0, cnt1, cnt2, p;
      bool cached = false;
#pragma unroll
      for (int k = 0; k < UPTS; ++k)
        if (c0 > 0 && slot == k) {
#pragma unroll
          for (int e = 0; e < WOFF; ++e) w[e] = e == 0 ? 1.0f : 0.0f;
#pragma unroll
          for (int e = 0; e < NDIM * NT; ++e) w[WOFF + e] = cw[k][e];
          rowoff = crow[k];
          cnt0 = ccnt[k] & 255; cnt1 = (ccnt[k] >> 8) & 255; cnt2 = (ccnt[k] >> 16) & 255;
          p = cpnt[k];
          cached = true;
        }
      if (!cached) {
        const int q = start + i;
        p = perm[q];
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
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            float x = (d < 3 - NDIM && k == 0) ? 1.0f : 0.0f;
            if (d >= 3 - NDIM && k < r[d].count)
              x = window(G.kernel,
                         fminf(fabsf((float)(r[d].g0 + k) - t[d]) * G.inv_half[d], 1.0f),
                         G.beta);
            w[d * NT + k] = x;
          }
        rowoff = (lo[0] * B.win[1] + lo[1]) * B.win[2] + lo[2];
        cnt0 = r[0].count; cnt1 = r[1].count; cnt2 = r[2].count;
#pragma unroll
        for (int k = 0; k < UPTS; ++k)
          if (slot == k) {
#pragma unroll
            for (int e = 0; e < NDIM * NT; ++e) cw[k][e] = w[WOFF + e];
            crow[k] = rowoff;
            ccnt[k] = cnt0 | (cnt1 << 8) | (cnt2 << 16);
            cpnt[k] = p;
          }
      }
      T acc[CC];
#pragma unroll
      for (int c = 0; c < CC; ++c) acc[c] = C::zero();
      // common case: four taps on every axis (all but points on integer
      // coordinates for widths in (3, 4]).  Branch-free, so each x row's
      // 16 x CC LDS reads issue together and their latency overlaps.
      const bool four = NT >= 4 && cnt1 == 4 && cnt2 == 4 && (NDIM < 3 || cnt0 == 4);
      if (NDIM >= 2 && four) {
#pragma unroll
        for (int kx = 0; kx < (NDIM >= 3 ? 4 : 1); ++kx) {
          const float wx = w[kx];
          if constexpr (ILV) {
            // RY y rows of the x row at a time: fewer registers in flight
            // (117 instead of 148 VGPRs with RY = 2, so 4 instead of 3
            // waves per SIMD; the radial workload's ungrid ran 1.59 ->
            // 1.36 ms), same summation order
            constexpr int RY = ILV_RY;
            const float4* rq = reinterpret_cast<const float4*>(lds) + rowoff + kx * ILV_W * ILV_W;
#pragma unroll
            for (int hy = 0; hy < 4; hy += RY) {
              float4 qv[RY][4];
#pragma unroll
              for (int ky = 0; ky < RY; ++ky)
#pragma unroll
                for (int kz = 0; kz < 4; ++kz) qv[ky][kz] = rq[(hy + ky) * ILV_W + kz];
#pragma unroll
              for (int ky = 0; ky < RY; ++ky) {
                const float wxy = wx * w[NT + hy + ky];
#pragma unroll
                for (int kz = 0; kz < 4; ++kz) {
                  const float wt = wxy * w[2 * NT + kz];
                  C::fma(acc[0], wt, make_float2(qv[ky][kz].x, qv[ky][kz].y));
                  C::fma(acc[CC - 1], wt, make_float2(qv[ky][kz].z, qv[ky][kz].w));
                }
              }
            }
            continue;
          }
          if constexpr (NDIM == 2 && CC >= 4) {
            // 2-D with 4+ channels: one y row of 4 x CC reads at a time
            // (the whole 4 x 4 x CC window would take 16 CC registers)
            const T* rp = lds + rowoff;
#pragma unroll
            for (int ky = 0; ky < 4; ++ky) {
              T tr[4][CC];
#pragma unroll
              for (int kz = 0; kz < 4; ++kz)
#pragma unroll
                for (int c = 0; c < CC; ++c) tr[kz][c] = rp[ky * B.win[2] + c * B.wcells + kz];
              const float wxy = wx * w[NT + ky];
#pragma unroll
              for (int kz = 0; kz < 4; ++kz) {
                const float wt = wxy * w[2 * NT + kz];
#pragma unroll
                for (int c = 0; c < CC; ++c) C::fma(acc[c], wt, tr[kz][c]);
              }
            }
            continue;
          }
          T tv[4][4][CC];
          {
            const T* rp = lds + rowoff + kx * B.win[1] * B.win[2];
#pragma unroll
            for (int ky = 0; ky < 4; ++ky)
#pragma unroll
              for (int kz = 0; kz < 4; ++kz)
#pragma unroll
                for (int c = 0; c < CC; ++c) tv[ky][kz][c] = rp[ky * B.win[2] + c * B.wcells + kz];
          }
#pragma unroll
          for (int ky = 0; ky < 4; ++ky) {
            const float wxy = wx * w[NT + ky];
#pragma unroll
            for (int kz = 0; kz < 4; ++kz) {
              const float wt = wxy * w[2 * NT + kz];
#pragma unroll
              for (int c = 0; c < CC; ++c) C::fma(acc[c], wt, tv[ky][kz][c]);
            }
          }
        }
      } else {
#pragma unroll 1
        for (int kx = 0; kx < cnt0; ++kx) {
          float wx = 1.0f;
#pragma unroll
          for (int k = 0; k < NT; ++k)
            if (k == kx) wx = w[k];
#pragma unroll
          for (int ky = 0; ky < NT; ++ky) {
            if (ky >= cnt1) break;
            const float wxy = wx * w[NT + ky];
            const int cell = rowoff + (kx * B.win[1] + ky) * B.win[2];
            const T* rowp = ILV ? lds + cell * CC : lds + cell;
#pragma unroll
            for (int kz = 0; kz < NT; ++kz) {
              if (kz >= cnt2) break;
              const float wt = wxy * w[2 * NT + kz];
#pragma unroll
              for (int c = 0; c < CC; ++c)
                C::fma(acc[c], wt, ILV ? rowp[kz * CC + c] : rowp[c * B.wcells + kz]);
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < CC; ++c)
        if (c < nc) o[pm ? p * nbatch + c0 + c : (c0 + c) * B.npts + p] = acc[c];
    }
  }
}

template <int NDIM, bool CPLX, int CC, int NT, bool ILV>
__global__ __launch_bounds__(BIN_THREADS) void ungrid_binned_kernel(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    InterpGeom G, BinGeom B, const int* __restrict__ perm,
    const float* __restrict__ slocs, const int4* __restrict__ sub,
    const int* __restrict__ nsub_ptr, bool pm) {
  // The grid is sized for the plan's subproblem capacity; remap over the
  // live count only, so that every XCD gets its share of the real work
  // (hardware hands workgroup b to XCD b % NUM_XCD).
  const int64_t nsub = *nsub_ptr;
  const int64_t n8 = (nsub + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  if ((int64_t)blockIdx.x >= n8) return;
  const int64_t s = xcd_remap(blockIdx.x, n8);
  if (s >= nsub) return;
  ungrid_sub<NDIM, CPLX, CC, NT, ILV>(vals, out, nbatch, G, B, perm, slocs, sub[s], pm);
}

// Point-major interpolation output (npts x nbatch, from ungrid_sub with pm)
// to the caller's channel-major layout (nbatch x npts): tiles of UNPK_P
// points through LDS, 8 channels at a time; the reads run along the
// point-major rows and the writes along each channel's points.
constexpr int UNPK_P = 128;

template <bool CPLX>
__global__ __launch_bounds__(256) void unpack_points_kernel(const float* __restrict__ src,
                                                            float* __restrict__ dst,
                                                            int64_t nbatch, int64_t npts) {
  using T = typename Ch<CPLX>::T;
  constexpr int TS = UNPK_P + 1;  // padded row: the 8 channels fall in different banks
  __shared__ T tile[8 * TS];
  const T* __restrict__ sv = reinterpret_cast<const T*>(src);
  T* __restrict__ dv = reinterpret_cast<T*>(dst);
  for (int64_t p0 = (int64_t)blockIdx.x * UNPK_P; p0 < npts;
       p0 += (int64_t)gridDim.x * UNPK_P) {
    const int np = (int)(npts - p0 < UNPK_P ? npts - p0 : UNPK_P);
    for (int64_t c0 = 0; c0 < nbatch; c0 += 8) {
      const int cc = (int)(nbatch - c0 < 8 ? nbatch - c0 : 8);
      __syncthreads();  // the previous tile's writes are done
      for (int e = threadIdx.x; e < np * cc; e += 256) {
        const int p = e / cc, c = e - p * cc;
        tile[c * TS + p] = sv[(p0 + p) * nbatch + c0 + c];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < cc * UNPK_P; e += 256) {
        const int c = e / UNPK_P, p = e - c * UNPK_P;
        if (p < np) dv[(c0 + c) * npts + p0 + p] = tile[c * TS + p];
      }
    }
  }
}

// Pre-pass for batched gridding: zero the output region of every split
// (shared) tile, once per tile (by its first item).  Called by one
// workgroup per item.
template <bool CPLX>
__device__ __forceinline__ void zero_shared_item(float* __restrict__ out,
                                                 int64_t nbatch,
                                                 const InterpGeom& G,
                                                 const BinGeom& B, int4 item,
                                                 const int* __restrict__ gstarts) {
  if (!item.w || item.y != gstarts[item.x]) return;
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
    out[(c * G.ncells + (gx * G.n[1] + gy) * G.n[2] + gz) * comp + part] = 0.0f;
  }
}

// Pre-pass for gridding of one problem: zero the output region of every
// split tile (more than g_cap touches, so its items add with atomics).  One
// thread per (tile, cell), cells z fastest so that a wave's stores cover
// whole rows; every channel.  Cells of unsplit tiles return at once.
// Split tiles cluster (the dense centre of a radial trajectory), so work
// is spread by cell rather than by tile.
template <bool CPLX>
__global__ __launch_bounds__(256) void grid_zero_split_kernel(
    float* __restrict__ out, int64_t nbatch, InterpGeom G, BinGeom B,
    const int* __restrict__ gcounts, int64_t ntiles, int g_cap) {
  using T = typename Ch<CPLX>::T;
  T* __restrict__ o = reinterpret_cast<T*>(out);
  const int tc = B.tcells, t12 = B.tile[1] * B.tile[2];
  const int64_t total = ntiles * tc;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = idx / tc;
    if (gcounts[t] <= g_cap) continue;
    const int cell = (int)(idx - t * tc);
    int origin[3];
    int r = (int)t;
    for (int d = 2; d >= 0; --d) {
      origin[d] = (r % B.ntiles[d]) * B.tile[d];
      r /= B.ntiles[d];
    }
    const int lx = cell / t12, rem = cell - lx * t12;
    const int ly = rem / B.tile[2], lz = rem - ly * B.tile[2];
    const int64_t gx = origin[0] + lx, gy = origin[1] + ly, gz = origin[2] + lz;
    if (gx >= G.n[0] || gy >= G.n[1] || gz >= G.n[2]) continue;
    const int64_t g = (gx * G.n[1] + gy) * G.n[2] + gz;
    for (int64_t c = 0; c < nbatch; ++c) o[c * G.ncells + g] = Ch<CPLX>::zero();
  }
}

// Launch the split-tile pre-pass for one problem's plan.
template <bool CPLX>
static int zero_split_tiles(float* out, int64_t nbatch, const InterpGeom& G, const BinGeom& B,
                            const tnl_plan_layout& L, const char* ws, hipStream_t s) {
  const int64_t nwg = (L.g_ntiles_total * B.tcells + 255) / 256;
  grid_zero_split_kernel<CPLX><<<(unsigned)(nwg < 16384 ? nwg : 16384), 256, 0, s>>>(
      out, nbatch, G, B, (const int*)(ws + L.off_gcounts), L.g_ntiles_total, (int)L.g_cap);
  TNL_LAUNCH_CHECK("grid_zero_split");
  return TNL_OK;
}

// Gridding, "lanes = taps".  A workgroup of NW waves owns one tile
// accumulator in LDS: NP channel planes (row and plane strides padded so the
// 64 taps of one instruction spread over the LDS banks), split
// evenly over the waves, so no two waves ever write the same word.  Touches
// are taken in batches of PB points, in three phases:
//   1. lanes = points, by role: one wave per spatial axis computes the
//      first tap, the in-tile tap mask and the 1-D window values; one wave
//      stores the channel values.  Right after, the raw inputs of the next
//      batch are loaded into registers, so their HBM latency overlaps the
//      rest of this batch.
//   2. all waves: the 64 tap weights of every point (products of the 1-D
//      windows, stored tap-major) and a per-point record: LDS offset of tap
//      (0,0,0), the 64-bit mask of live taps and flags.
//   3. lanes = taps: every wave walks the batch point by point over its own
//      planes.  A lane holds one tap of a 4x4x4 block (3-D) or of an 8x8
//      block (2-D); it is live when its bit is set in the point's mask, and
//      its cell is the point's offset plus a per-lane constant.  Within an
//      instruction every lane addresses a distinct cell and only one wave
//      owns a plane, so the accumulation is a plain read-modify-write (LDS
//      float atomics run about one lane per clock).  Two rare cases take a
//      general path: points on integer coordinates in 3-D (a fifth tap,
//      covered by a second pass over the 61 taps with an index 4) and
//      footprints that wrap around the grid inside the tile.
// The tile then goes to HBM once, coalesced along z.
constexpr int GW = 64;  // wave size
constexpr int PB = 32;  // points per batch

constexpr int GF_SHELL = 1;  // some axis has a live fifth tap (3-D)
constexpr int GF_WRAP = 2;   // the footprint wraps around the grid in this tile

// LDS carve-up of the gridding kernel (host and device agree on it).
struct GridLds {
  int off_dim;   // int2 [3][PB]: {lo * stride, mask | wrap << 8 | lo << 16}
  int off_w1;    // float [PB][w1s]: 1-D window values
  int off_v;     // float [PB][NP]: channel values
  int off_w64;   // float [64][PB + 1]: tap weights, tap-major
  int off_meta;  // int4 [PB]: {offset of tap (0,0,0), mask lo, mask hi, flags}
  int w1s;
  int bytes;
};

__host__ __device__ inline int align16(int x) { return (x + 15) / 16 * 16; }

__host__ __device__ inline GridLds grid_lds_layout(const BinGeom& B, int planes, int nt) {
  GridLds g;
  const int pcells = B.tile[0] * B.gxs;
  g.w1s = (3 * nt) | 1;  // odd row stride: conflict-free per-point access
  g.off_dim = align16(planes * pcells * 4);
  g.off_w1 = g.off_dim + 3 * PB * 8;
  g.off_v = align16(g.off_w1 + PB * g.w1s * 4);
  g.off_w64 = g.off_v + PB * planes * 4;
  g.off_meta = align16(g.off_w64 + GW * (PB + 1) * 4);
  g.bytes = g.off_meta + PB * 16;
  return g;
}

template <int N>
struct VecF;
template <> struct VecF<1> { using T = float; };
template <> struct VecF<2> { using T = float2; };
template <> struct VecF<4> { using T = float4; };

// This lane's tap in the main block (3-D: 4x4x4, 2-D: 8x8) and, in 3-D,
// its tap of the second pass over the 61 taps with an index 4.
struct GridLanes {
  int mk[3], sk[3];
  bool shell_ok;
};

template <int NDIM>
__device__ __forceinline__ GridLanes make_grid_lanes(int lane) {
  GridLanes g;
  g.shell_ok = false;
  g.sk[0] = g.sk[1] = g.sk[2] = 0;
  if (NDIM >= 3) {
    g.mk[0] = lane >> 4; g.mk[1] = (lane >> 2) & 3; g.mk[2] = lane & 3;
    int seen = 0;
    for (int a = 0; a < 5; ++a)
      for (int b2 = 0; b2 < 5; ++b2)
        for (int c2 = 0; c2 < 5; ++c2)
          if (a == 4 || b2 == 4 || c2 == 4) {
            if (seen == lane) { g.sk[0] = a; g.sk[1] = b2; g.sk[2] = c2; g.shell_ok = true; }
            ++seen;
          }
  } else {
    g.mk[0] = 0; g.mk[1] = lane >> 3; g.mk[2] = lane & 7;
  }
  return g;
}

// One work item (a tile and the touches listed for it) of one problem,
// every channel chunk.  Called by all threads of the workgroup.
template <int NDIM, bool CPLX, int CC, int NT, int NW>
__device__ __forceinline__ void grid_item(
    const GridLanes& ln, const float* __restrict__ vals, float* __restrict__ out,
    int64_t nbatch, const InterpGeom& G, const BinGeom& B,
    const int* __restrict__ perm, const float* __restrict__ slocs,
    const int* __restrict__ tlist, int4 item) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  constexpr int COMP = CPLX ? 2 : 1;
  constexpr int NP = CC * COMP;  // accumulator planes
  constexpr int PPW = NP / NW;   // planes per wave
  constexpr int NR = 4 / NW;     // phase-1 roles per wave
  constexpr int NTH = GW * NW;
  constexpr int VW = PPW % 4 == 0 ? 4 : (PPW % 2 == 0 ? 2 : 1);
  constexpr int W64S = PB + 1;  // tap-major weight row stride
  static_assert(NP % NW == 0 && 4 % NW == 0, "planes and roles split evenly");
  static_assert(NT <= 8, "tap masks hold 8 bits per axis");
  static_assert(NTH % PB == 0, "phase 2 maps threads to (point, tap group)");
  using V = typename VecF<VW>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int pcells = B.tile[0] * B.gxs;  // cells per plane
  const GridLds LG = grid_lds_layout(B, NP, NT);
  float* acc = reinterpret_cast<float*>(lds_raw);  // [NP][pcells]
  int2* s_dim = reinterpret_cast<int2*>(lds_raw + LG.off_dim);
  float* s_w1 = reinterpret_cast<float*>(lds_raw + LG.off_w1);
  float* s_v = reinterpret_cast<float*>(lds_raw + LG.off_v);
  float* s_w64 = reinterpret_cast<float*>(lds_raw + LG.off_w64);
  int4* s_meta = reinterpret_cast<int4*>(lds_raw + LG.off_meta);
  const int w1s = LG.w1s;
  const int tid = threadIdx.x;
  const int lane = tid & (GW - 1);
  const int wv = __builtin_amdgcn_readfirstlane(tid / GW);
  const T* __restrict__ v = reinterpret_cast<const T*>(vals);
  const int dstride[3] = {B.gxs, B.gys, 1};
  const int mk[3] = {ln.mk[0], ln.mk[1], ln.mk[2]};
  const int sk[3] = {ln.sk[0], ln.sk[1], ln.sk[2]};
  const bool shell_ok = ln.shell_ok;
  const int mkoff = mk[0] * dstride[0] + mk[1] * dstride[1] + mk[2];
  float* __restrict__ accw = acc + wv * PPW * pcells;  // this wave's planes
  const float* __restrict__ w64l = s_w64 + lane * W64S;

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
    // raw inputs of one batch, per role: a coordinate or the values
    float pt[NR];
    T pv[CC];
    auto load = [&](int b0, int nb) {
      if (lane < nb) {
        const int q = tlist[start + b0 + lane];
#pragma unroll
        for (int sr = 0; sr < NR; ++sr) {
          const int role = wv + sr * NW;
          if (role < 3) {
            pt[sr] = role >= 3 - NDIM ? slocs[(int64_t)q * NDIM + role - (3 - NDIM)] : 0.f;
          } else {
            const int p = perm[q];
#pragma unroll
            for (int c = 0; c < CC; ++c)
              pv[c] = c < nc ? v[(c0 + c) * B.npts + p] : C::zero();
          }
        }
      }
    };
    {
      float4* a4 = reinterpret_cast<float4*>(acc);
      const int n4 = NP * pcells / 4;
      for (int i = tid; i < n4; i += NTH) a4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int i = n4 * 4 + tid; i < NP * pcells; i += NTH) acc[i] = 0.f;
    }
    load(0, count < PB ? count : PB);
    for (int b0 = 0; b0 < count; b0 += PB) {
      const int nb = count - b0 < PB ? count - b0 : PB;
      __syncthreads();
      // ---- phase 1, lanes = points: per-axis records and values
      if (lane < nb) {
#pragma unroll
        for (int sr = 0; sr < NR; ++sr) {
          const int role = wv + sr * NW;
          if (role < 3) {
            const int d = role;
            int lo = 0, cnt = 1;
            float t = 0.f;
            int64_t g0 = 0;
            if (d >= 3 - NDIM) {
              t = pt[sr];
              const TapRange r = tap_range(t, G.half[d], G.ntap[d], G.n[d], G.pad);
              const int64_t gw =
                  G.pad == TNL_PAD_CIRCULAR ? mod_pos(r.g0, G.n[d]) : r.g0;
              lo = (int)(gw - origin[d]);
              cnt = r.count;
              g0 = r.g0;
            }
            // in-tile mask over taps k < cnt, and whether a tap wraps
            // past the end of the grid into this tile
            int m = 0, wrap = 0;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              if (k < cnt) {
                int l = lo + k;
                if (G.pad == TNL_PAD_CIRCULAR && origin[d] + l >= G.n[d]) {
                  l -= (int)G.n[d];
                  wrap = 1;
                }
                if (l >= 0 && l < B.tile[d]) m |= 1 << k;
              }
            }
            s_dim[d * PB + lane] =
                make_int2(lo * dstride[d], m | (wrap << 8) | (int)((unsigned)lo << 16));
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              float w = (d < 3 - NDIM && k == 0) ? 1.0f : 0.0f;
              if (d >= 3 - NDIM && k < cnt)
                w = window(G.kernel,
                           fminf(fabsf((float)(g0 + k) - t) * G.inv_half[d], 1.0f),
                           G.beta);
              s_w1[lane * w1s + d * NT + k] = w;
            }
          } else {
            V* dst = reinterpret_cast<V*>(s_v + lane * NP);
            float tmp[NP];
#pragma unroll
            for (int c = 0; c < CC; ++c) {
              if constexpr (CPLX) {
                tmp[2 * c] = pv[c].x;
                tmp[2 * c + 1] = pv[c].y;
              } else {
                tmp[c] = pv[c];
              }
            }
#pragma unroll
            for (int i = 0; i < NP / VW; ++i) {
              V x;
              if constexpr (VW == 4) x = make_float4(tmp[4 * i], tmp[4 * i + 1], tmp[4 * i + 2], tmp[4 * i + 3]);
              else if constexpr (VW == 2) x = make_float2(tmp[2 * i], tmp[2 * i + 1]);
              else x = tmp[i];
              dst[i] = x;
            }
          }
        }
      }
      if (b0 + PB < count) load(b0 + PB, count - b0 - PB < PB ? count - b0 - PB : PB);
      __syncthreads();
      // ---- phase 2: tap weights (tap-major) and point records
      {
        const int j = tid % PB;
        if (j < nb) {
          const float* w1 = s_w1 + j * w1s;
          // 8 groups of 8 taps: 3-D group g = (k0 = g / 2, k1 in 2(g % 2) + {0, 1},
          // k2 in 0..3); 2-D group g = (k1 = g, k2 in 0..7)
          for (int g = tid / PB; g < 8; g += NTH / PB) {
            float* dst = s_w64 + (g * 8) * W64S + j;
            if (NDIM >= 3) {
              const float wx = w1[g >> 1];
              const float wxy0 = wx * w1[NT + 2 * (g & 1)];
              const float wxy1 = wx * w1[NT + 2 * (g & 1) + 1];
#pragma unroll
              for (int k2 = 0; k2 < 4; ++k2) {
                const float wz = w1[2 * NT + k2];
                dst[k2 * W64S] = wxy0 * wz;
                dst[(4 + k2) * W64S] = wxy1 * wz;
              }
            } else {
              const float wy = w1[NT + g];
#pragma unroll
              for (int k2 = 0; k2 < 8; ++k2)
                dst[k2 * W64S] = k2 < NT ? wy * w1[2 * NT + k2] : 0.0f;
            }
          }
          if (tid < PB) {
            const int2 a0 = s_dim[j], a1 = s_dim[PB + j], a2 = s_dim[2 * PB + j];
            const unsigned m0 = a0.y & 255, m1 = a1.y & 255, m2 = a2.y & 255;
            uint64_t m64 = 0;
            if (NDIM >= 3) {
              unsigned m12 = 0;
#pragma unroll
              for (int k1 = 0; k1 < 4; ++k1)
                if ((m1 >> k1) & 1) m12 |= (m2 & 15u) << (4 * k1);
#pragma unroll
              for (int k0 = 0; k0 < 4; ++k0)
                if ((m0 >> k0) & 1) m64 |= (uint64_t)m12 << (16 * k0);
            } else {
#pragma unroll
              for (int k1 = 0; k1 < 8; ++k1)
                if ((m1 >> k1) & 1) m64 |= (uint64_t)m2 << (8 * k1);
            }
            const int mall = a0.y | a1.y | a2.y;
            const int flags = (NDIM >= 3 && (mall & 16) ? GF_SHELL : 0) |
                              (mall & 256 ? GF_WRAP : 0);
            s_meta[j] = make_int4(a0.x + a1.x + a2.x, (int)(unsigned)m64,
                                  (int)(unsigned)(m64 >> 32), flags);
          }
        }
      }
      __syncthreads();
      // ---- phase 3, lanes = taps: this wave's planes, point by point.
      // The next point's record, values and weight are loaded during this
      // point's read-modify-write (one wave's LDS operations complete in
      // order, so its reads see its earlier writes).
      const int nbu = __builtin_amdgcn_readfirstlane(nb);
      int4 n_meta = s_meta[0];
      float n_w = w64l[0];
      V n_v[PPW / VW];
      {
        const V* src = reinterpret_cast<const V*>(s_v + wv * PPW);
#pragma unroll
        for (int i = 0; i < PPW / VW; ++i) n_v[i] = src[i];
      }
      for (int j = 0; j < nbu; ++j) {
        const int4 meta = n_meta;
        const float w = n_w;
        float vv[PPW];
#pragma unroll
        for (int i = 0; i < PPW / VW; ++i) {
          const V x = n_v[i];
          if constexpr (VW == 4) {
            vv[4 * i] = x.x; vv[4 * i + 1] = x.y; vv[4 * i + 2] = x.z; vv[4 * i + 3] = x.w;
          } else if constexpr (VW == 2) {
            vv[2 * i] = x.x; vv[2 * i + 1] = x.y;
          } else {
            vv[i] = x;
          }
        }
        if (j + 1 < nbu) {
          n_meta = s_meta[j + 1];
          n_w = w64l[j + 1];
          const V* src = reinterpret_cast<const V*>(s_v + (j + 1) * NP + wv * PPW);
#pragma unroll
          for (int i = 0; i < PPW / VW; ++i) n_v[i] = src[i];
        }
        const unsigned mlo = __builtin_amdgcn_readfirstlane(meta.y);
        const unsigned mhi = __builtin_amdgcn_readfirstlane(meta.z);
        const int flags = __builtin_amdgcn_readfirstlane(meta.w);
        const uint64_t m64 = ((uint64_t)mhi << 32) | mlo;
        if ((m64 >> lane) & 1) {
          int cell = meta.x + mkoff;
          if (flags & GF_WRAP) {  // footprint wraps around the grid in this tile
            cell = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              int l = (s_dim[d * PB + j].y >> 16) + mk[d];
              if (origin[d] + l >= G.n[d]) l -= (int)G.n[d];
              cell += l * dstride[d];
            }
          }
          float* cp = accw + cell;
          float old[PPW];
#pragma unroll
          for (int c = 0; c < PPW; ++c) old[c] = cp[c * pcells];
#pragma unroll
          for (int c = 0; c < PPW; ++c) cp[c * pcells] = fmaf(w, vv[c], old[c]);
        }
        if (NDIM >= 3 && (flags & GF_SHELL)) {
          // second pass: taps with some index 4 (points on integer coordinates)
          const int2 a0 = s_dim[j], a1 = s_dim[PB + j], a2 = s_dim[2 * PB + j];
          const bool ok = shell_ok && ((a0.y >> sk[0]) & (a1.y >> sk[1]) & (a2.y >> sk[2]) & 1);
          if (ok) {
            const int lo3[3] = {a0.y >> 16, a1.y >> 16, a2.y >> 16};
            int cell = 0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              int l = lo3[d] + sk[d];
              if (G.pad == TNL_PAD_CIRCULAR && origin[d] + l >= G.n[d]) l -= (int)G.n[d];
              cell += l * dstride[d];
            }
            const float* w1 = s_w1 + __builtin_amdgcn_readfirstlane(j * w1s);
            const float ws3 = w1[sk[0]] * w1[NT + sk[1]] * w1[2 * NT + sk[2]];
            float* cp = accw + cell;
            float old[PPW];
#pragma unroll
            for (int c = 0; c < PPW; ++c) old[c] = cp[c * pcells];
#pragma unroll
            for (int c = 0; c < PPW; ++c) cp[c * pcells] = fmaf(ws3, vv[c], old[c]);
          }
        }
      }
    }
    __syncthreads();
    // ---- write the tile: coalesced along z, one channel at a time
    T* __restrict__ o = reinterpret_cast<T*>(out);
    const int t12 = B.tile[1] * B.tile[2];
    for (int cell = tid; cell < B.tcells; cell += NTH) {
      const int lx = cell / t12, rem = cell - lx * t12;
      const int ly = rem / B.tile[2], lz = rem - ly * B.tile[2];
      const int64_t gx = origin[0] + lx, gy = origin[1] + ly,
                    gz = origin[2] + lz;
      if (gx >= G.n[0] || gy >= G.n[1] || gz >= G.n[2]) continue;
      const int64_t g0 = (gx * G.n[1] + gy) * G.n[2] + gz;
      const int pc = lx * dstride[0] + ly * dstride[1] + lz;
#pragma unroll
      for (int c = 0; c < CC; ++c) {
        if (c >= nc) break;
        const int64_t gi = (c0 + c) * G.ncells + g0;
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
    }
    __syncthreads();
  }
}

template <int NDIM, bool CPLX, int CC, int NT, int NW>
__global__ __launch_bounds__(GW * NW) void grid_binned_kernel(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    InterpGeom G, BinGeom B, const int* __restrict__ perm,
    const float* __restrict__ slocs, const int* __restrict__ tlist,
    const int4* __restrict__ items, const int* __restrict__ nitems_ptr) {
  const GridLanes ln = make_grid_lanes<NDIM>(threadIdx.x & (GW - 1));
  const int nitems = *nitems_ptr;
  for (int64_t it = xcd_remap(blockIdx.x, gridDim.x); it < nitems;
       it += gridDim.x)
    grid_item<NDIM, CPLX, CC, NT, NW>(ln, vals, out, nbatch, G, B, perm, slocs,
                                      tlist, items[it]);
}

// ---------------------------------------------------------------------------
// Gridding with matrix-core accumulation (3-D 8x8x8 and 2-D 16x32 tiles;
// 512 cells either way).  Described here for 3-D; the 2-D layout follows
// (MfShape).
//
// For one tile, out[cell][plane] = sum_p A[cell][p] V[p][plane] with
// A[cell][p] = wx_p(x) wy_p(y) wz_p(z), the point's separable weight at the
// cell (zero off its footprint), and V the point's channel values (planes:
// re/im of up to 8 coils, or up to 16 real channels).  That is a
// (512 x P) by (P x 16) product.  v_mfma_f32_16x16x4f32 takes a block of 16
// cells (one x plane, a 4x4 (y, z) patch) by 16 planes by 4 points; lane l
// holds A[l % 16][l / 16], B[l / 16][l % 16] and D[4 (l / 16) + i][l % 16].
// A block that none of the points of a pair of such groups touches is
// skipped (a wave-uniform mask), so the work follows the footprints rather
// than the tile volume.
//
// Workgroup: 4 waves on one tile.  Waves 0, 1 hold the even x planes (16
// blocks, 64 accumulator registers per lane), waves 2, 3 the odd ones, so a
// footprint (4 or 5 consecutive planes) gives both pairs about the same
// work; the two waves of a pair take alternate pairs of point groups.
// Touches are taken in batches of MB points (one per thread):
//   1. lanes = points: first taps, the 8-cell tile-local window of every
//      axis (zero off the footprint; wrapped taps land where they wrap to),
//      the channel values, and the mask of touched blocks OR-ed over each
//      group of 4 points.  The next batch's raw inputs (of this tile, or of
//      the workgroup's next tile or channel chunk) are loaded right after.
//   2. waves: per pair of groups of 4 points, A = products of the windows,
//      B = the values, two MFMAs per live block.
// At the end of the tile the two partial sums of each x plane meet in LDS
// (one wave stores, the other adds) and the tile goes to HBM once,
// coalesced along z.
typedef float mf32x4 __attribute__((ext_vector_type(4)));

constexpr int MF_THREADS = 256;
constexpr int MF_PSTR = 516;  // staging plane stride (floats): 512 + 4
constexpr int MF_NPL = 16;    // accumulator planes

// Per-dimension shape of the matrix-core gridding kernel.  A tile has 512
// cells in 32 blocks of 16; block b = 4 u + v, where u in [0, 8) is split
// between the wave pairs by parity and v in [0, 4) is not:
//   3-D, 8x8x8 tiles: u = x plane, v = 4x4 (y, z) patch (2 (y / 4) + z / 4)
//   2-D, 16x32 tiles: u = z / 4, v = y / 4
// Per point of a batch, LDS holds the "U" window (3-D: x, 8 cells; 2-D: z,
// 32 cells), the "V" windows (3-D: y then z, 8 cells each; 2-D: y, 16
// cells) and the 16 channel planes.
template <int NDIM>
struct MfShape;
template <>
struct MfShape<3> {
  static constexpr int MB = 256;  // points per batch (one per thread)
  static constexpr int RA = 8;    // floats per point: U window
  static constexpr int RB = 16;   // floats per point: V windows
};
template <>
struct MfShape<2> {
  static constexpr int MB = 128;
  static constexpr int RA = 32;
  static constexpr int RB = 16;
};

template <int NDIM>
constexpr size_t mf_lds_bytes() {
  using S = MfShape<NDIM>;
  const size_t batch = (size_t)S::MB * (S::RA + S::RB + MF_NPL) * 4 + (size_t)S::MB;
  const size_t stage = (size_t)MF_NPL * MF_PSTR * 4;
  return batch > stage ? batch : stage;
}

// Slot orders of the tile-local windows, chosen so that a lane reads the
// four values it needs with one b128 load (or a pair with one b64 load).
enum MfOrder {
  MF_X8,    // 3-D x: [W0, W2, W4, W6, W1, W3, W5, W7] (a wave's parity)
  MF_PAIR8, // 3-D y, z: [W0, W4, W1, W5, W2, W6, W3, W7] (pairs {j, j + 4})
  MF_Y16,   // 2-D y: slot 4 (y % 4) + y / 4 (lane y % 4 reads y / 4 = 0..3)
  MF_Z32    // 2-D z: slot 16 ((z / 4) % 2) + 4 (z % 4) + z / 8 (parity, lane)
};

template <int ORD>
__device__ __forceinline__ int mf_slot(int l) {
  if constexpr (ORD == MF_X8) return (l & 1) * 4 + (l >> 1);
  else if constexpr (ORD == MF_PAIR8) return 2 * (l & 3) + (l >> 2);
  else if constexpr (ORD == MF_Y16) return (l & 3) * 4 + (l >> 2);
  else return ((l >> 2) & 1) * 16 + (l & 3) * 4 + (l >> 3);
}

// TL-cell tile-local window of axis d for the point at coordinate t,
// written to row[0..TL), which the caller has cleared: each tap's weight
// lands in the slot of the cell it falls on (taps wrap around the grid
// under circular padding; no two taps share a cell, as the footprint never
// exceeds the grid).  Returns the mask of cells some tap falls on.
template <int TL, int ORD>
__device__ __forceinline__ unsigned mf_axis_place(float t, int d, int origin,
                                                  const InterpGeom& G,
                                                  float* __restrict__ row) {
  unsigned pm = 0;
  const TapRange r = tap_range(t, G.half[d], G.ntap[d], G.n[d], G.pad);
  // The first kept tap wrapped into [0, n) under circular padding: one add
  // or subtract for coordinates within a grid length of the grid, a 64-bit
  // modulo only beyond.  Cell indices are 32-bit from here (grid axes are
  // far below 2^31), and tap k sits at g0f + k, exactly (float)(g0 + k)
  // while |g0| < 2^24.
  const int n = (int)G.n[d];
  int64_t a = r.g0;
  if (G.pad == TNL_PAD_CIRCULAR) {
    if (a < 0) a += n;
    else if (a >= n) a -= n;
    if (a < 0 || a >= n) a = mod_pos(r.g0, G.n[d]);
  }
  const int lo = (int)a - origin;
  const float g0f = (float)r.g0;
  // not unrolled: one copy of the window code per axis keeps the kernel
  // small in the instruction cache
#pragma unroll 1
  for (int k = 0; k < r.count; ++k) {
    int l = lo + k;
    if (G.pad == TNL_PAD_CIRCULAR && origin + l >= n) l -= n;
    if ((unsigned)l < (unsigned)TL) {
      row[mf_slot<ORD>(l)] =
          window(G.kernel, fminf(fabsf((g0f + (float)k) - t) * G.inv_half[d], 1.0f),
                 G.beta);
      pm |= 1u << l;
    }
  }
  return pm;
}

// Point-major record of one point and channel chunk for the matrix-core
// gridding kernel: the chunk's 16 planes (re/im of 8 coils, or 16 real
// channels; zero past nbatch), 64 bytes, in the caller's point order (a
// plain transpose of the coil-major input).  A workgroup moves a tile of
// 256 points through LDS: each channel row is read with consecutive
// lanes on consecutive points, and the tile's 16 KB of records is written
// with consecutive lanes on consecutive 16-byte words.
constexpr int MF_REC4 = 4;  // float4 per record
constexpr int PACK_PTS = 256;
constexpr int PACK_STR = MF_NPL + 1;  // LDS floats per point (odd: no bank conflicts)

template <bool CPLX>
__global__ __launch_bounds__(256) void pack_points_kernel(const float* __restrict__ vals,
                                                          int64_t npts, int64_t nbatch,
                                                          float4* __restrict__ rec) {
  constexpr int CH = CPLX ? MF_NPL / 2 : MF_NPL;
  __shared__ float tile[PACK_PTS * PACK_STR];
  const int64_t nch = (nbatch + CH - 1) / CH;
  const int64_t ntile = (npts + PACK_PTS - 1) / PACK_PTS;
  const int tid = threadIdx.x;
  for (int64_t u = blockIdx.x; u < ntile * nch; u += gridDim.x) {
    const int64_t ch = u / ntile;
    const int64_t p0 = (u - ch * ntile) * PACK_PTS;
    const int np = (int)(npts - p0 < PACK_PTS ? npts - p0 : PACK_PTS);
    __syncthreads();  // previous tile's records are written out
    if (tid < np) {
      const int64_t p = p0 + tid;
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int64_t cg = ch * CH + c;
        if constexpr (CPLX) {
          const float2 x = cg < nbatch ? reinterpret_cast<const float2*>(vals)[cg * npts + p]
                                       : make_float2(0.f, 0.f);
          tile[tid * PACK_STR + 2 * c] = x.x;
          tile[tid * PACK_STR + 2 * c + 1] = x.y;
        } else {
          tile[tid * PACK_STR + c] = cg < nbatch ? vals[cg * npts + p] : 0.f;
        }
      }
    }
    __syncthreads();
    float4* r = rec + (ch * npts + p0) * MF_REC4;
    for (int j = tid; j < np * MF_REC4; j += 256) {
      const float* t = tile + (j >> 2) * PACK_STR + 4 * (j & 3);
      r[j] = make_float4(t[0], t[1], t[2], t[3]);
    }
  }
}

// Shared-memory carve-up and lane roles of the matrix-core gridding
// kernels (one tile per workgroup of 4 waves).
template <int NDIM>
struct MfLds {
  float* wa;     // [MB][RA]: U window of each point of the batch
  float* wb;     // [MB][RB]: V windows
  float* v;      // [MB][16]: channel planes
  unsigned* km;  // [MB / 4]: block masks of the groups of 4 points
  float* stage;  // [16][MF_PSTR]: the tile's planes, after the batches
  __device__ explicit MfLds(unsigned char* raw) {
    using S = MfShape<NDIM>;
    wa = reinterpret_cast<float*>(raw);
    wb = wa + S::MB * S::RA;
    v = wb + S::MB * S::RB;
    km = reinterpret_cast<unsigned*>(v + S::MB * MF_NPL);
    stage = reinterpret_cast<float*>(raw);
  }
};

struct MfLane {
  int tid, lane, xpar, par;  // thread, lane, U parity of the wave, group parity
  int n16, k4, yy, zz;       // B/D column, A column / D row group, A row 4 yy + zz
  __device__ MfLane() {
    tid = threadIdx.x;
    lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    xpar = wv >> 1;
    par = wv & 1;
    n16 = lane & 15;
    k4 = lane >> 4;
    yy = n16 >> 2;
    zz = n16 & 3;
  }
};

// Phase 1, lanes = points: the tile-local windows of this thread's point
// (zero past nb), its channel planes, and the block masks OR-ed over each
// group of 4 points.
template <int NDIM, bool CPLX>
__device__ __forceinline__ void mf_phase1(const MfLds<NDIM>& S, const MfLane& ln, int nb,
                                          const float* pt, const typename Ch<CPLX>::T* pv,
                                          const int* origin, const InterpGeom& G) {
  using Sh = MfShape<NDIM>;
  constexpr int CH = CPLX ? MF_NPL / 2 : MF_NPL;
  const int tid = ln.tid;
  if (tid >= Sh::MB) return;
  const bool live = tid < nb;
  unsigned m = 0;
  float* ra = S.wa + tid * Sh::RA;
  float* rb = S.wb + tid * Sh::RB;
  // clear the window rows of this wave's 64 points (live or not: a dead
  // point's rows still meet zero values in the MFMAs), lane l taking the
  // 16-byte units l, l + 64, ... of the wave's contiguous rows, so each
  // store covers 1 KB without bank conflicts (a lane clearing its own row
  // put lanes 128 or 64 bytes apart on the same banks); the rows are this
  // wave's own, and its LDS stores complete in order, so no barrier
  {
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const int w0 = tid & ~63, lane = tid & 63;
    float4* a4 = reinterpret_cast<float4*>(S.wa + w0 * Sh::RA);
    float4* b4 = reinterpret_cast<float4*>(S.wb + w0 * Sh::RB);
#pragma unroll
    for (int i = 0; i < Sh::RA / 4; ++i) a4[64 * i + lane] = z4;
#pragma unroll
    for (int i = 0; i < Sh::RB / 4; ++i) b4[64 * i + lane] = z4;
  }
  if (live) {
    if constexpr (NDIM == 3) {
      const unsigned px = mf_axis_place<8, MF_X8>(pt[0], 0, origin[0], G, ra);
      const unsigned py = mf_axis_place<8, MF_PAIR8>(pt[1], 1, origin[1], G, rb);
      const unsigned pz = mf_axis_place<8, MF_PAIR8>(pt[2], 2, origin[2], G, rb + 8);
      m = tile_blocks32(px, py, pz);
    } else {
      const unsigned py = mf_axis_place<16, MF_Y16>(pt[0], 1, origin[1], G, rb);
      const unsigned pz = mf_axis_place<32, MF_Z32>(pt[1], 2, origin[2], G, ra);
      m = tile_blocks32_2d(py, pz);
    }
  }
  float tv[MF_NPL];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    if constexpr (CPLX) {
      tv[2 * c] = live ? pv[c].x : 0.f;
      tv[2 * c + 1] = live ? pv[c].y : 0.f;
    } else {
      tv[c] = live ? pv[c] : 0.f;
    }
  }
  float4* dv = reinterpret_cast<float4*>(S.v + tid * MF_NPL);
#pragma unroll
  for (int i = 0; i < MF_NPL / 4; ++i)
    dv[i] = make_float4(tv[4 * i], tv[4 * i + 1], tv[4 * i + 2], tv[4 * i + 3]);
  // blocks touched by the group of 4 points (bit 4 u + v)
  m |= (unsigned)__shfl_xor((int)m, 1);
  m |= (unsigned)__shfl_xor((int)m, 2);
  if ((tid & 3) == 0) S.km[tid >> 2] = m;
}

// Phase 2: this wave's blocks (U of one parity), every other pair of point
// groups (8 points; a pair shares one decision per block and gets two
// MFMAs per live block).  Lane i holds this wave's block mask of pair
// par + 2 i, OR-ed over its two groups; the pairs with live blocks are
// walked from a ballot, and the next pair's operands are read from LDS
// while the current pair's MFMAs run.
template <int NDIM>
__device__ __forceinline__ void mf_phase2(mf32x4 (&acc)[16], const MfLds<NDIM>& S,
                                          const MfLane& ln, int nb) {
  using Sh = MfShape<NDIM>;
  constexpr int RA = Sh::RA, RB = Sh::RB;
  const int lane = ln.lane, xpar = ln.xpar, par = ln.par;
  const int k4 = ln.k4, yy = ln.yy, zz = ln.zz, n16 = ln.n16;
  const int nkg = (nb + 3) >> 2;
  const int npg = (nkg + 1) >> 1;
  const int pgl = par + 2 * lane;
  unsigned lm = 0;
  if (lane < Sh::MB / 16 && pgl < npg) {
    unsigned mm = S.km[2 * pgl];
    if (2 * pgl + 1 < nkg) mm |= S.km[2 * pgl + 1];
    // this wave's blocks: u = xpar, xpar + 2, ...
#pragma unroll
    for (int j = 0; j < 4; ++j) lm |= ((mm >> (4 * (2 * j + xpar))) & 15u) << (4 * j);
  }
  uint64_t live = __ballot(lm != 0);
  struct Ops {
    float4 a4;  // U weights of this wave's 4 blocks u = 2 j + xpar
    float4 b4;  // 3-D: {wy(yy), wy(yy + 4), wz(zz), wz(zz + 4)}; 2-D: wy(4 v + yy)
    float bv;
  };
  // A (window) and B (value) operand parts of group g for this lane
  auto fetch = [&](int g) {
    const int pnt = g * 4 + k4;
    Ops o;
    if constexpr (NDIM == 3) {
      o.a4 = *reinterpret_cast<const float4*>(S.wa + pnt * RA + 4 * xpar);
      const float2 y2 = *reinterpret_cast<const float2*>(S.wb + pnt * RB + 2 * yy);
      const float2 z2 = *reinterpret_cast<const float2*>(S.wb + pnt * RB + 8 + 2 * zz);
      o.b4 = make_float4(y2.x, y2.y, z2.x, z2.y);
    } else {
      o.a4 = *reinterpret_cast<const float4*>(S.wa + pnt * RA + 16 * xpar + 4 * zz);
      o.b4 = *reinterpret_cast<const float4*>(S.wb + pnt * RB + 4 * yy);
    }
    o.bv = S.v[pnt * MF_NPL + n16];
    return o;
  };
  auto vweights = [&](const Ops& o, float* w4) {
    if constexpr (NDIM == 3) {
      w4[0] = o.b4.x * o.b4.z; w4[1] = o.b4.x * o.b4.w;
      w4[2] = o.b4.y * o.b4.z; w4[3] = o.b4.y * o.b4.w;
    } else {
      w4[0] = o.b4.x; w4[1] = o.b4.y; w4[2] = o.b4.z; w4[3] = o.b4.w;
    }
  };
  auto pair = [&](const Ops& o0, const Ops& o1, unsigned hm) {
    float wv0[4], wv1[4];
    vweights(o0, wv0);
    vweights(o1, wv1);
    const float wu0[4] = {o0.a4.x, o0.a4.y, o0.a4.z, o0.a4.w};
    const float wu1[4] = {o1.a4.x, o1.a4.y, o1.a4.z, o1.a4.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((hm >> (4 * j + q)) & 1u) {
          acc[4 * j + q] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              wu0[j] * wv0[q], o0.bv, acc[4 * j + q], 0, 0, 0);
          acc[4 * j + q] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              wu1[j] * wv1[q], o1.bv, acc[4 * j + q], 0, 0, 0);
        }
  };
  auto next = [&](int prev, unsigned* hm) {
    const bool more = live != 0;
    const int j = more ? __builtin_ctzll(live) : prev;
    if (more) live &= live - 1;
    *hm = more ? (unsigned)__builtin_amdgcn_readlane((int)lm, j) : 0u;
    return j;
  };
  // two operand sets in turn: while one pair's MFMAs issue, the other set
  // is being filled (no register copies between them)
  if (live) {
    unsigned ha, hb;
    int ia = next(0, &ha);
    Ops a0 = fetch(2 * (par + 2 * ia)), a1 = fetch(2 * (par + 2 * ia) + 1);
    for (;;) {
      const int ib = next(ia, &hb);
      const Ops b0 = fetch(2 * (par + 2 * ib)), b1 = fetch(2 * (par + 2 * ib) + 1);
      pair(a0, a1, ha);
      if (!hb) break;
      ia = next(ib, &ha);
      a0 = fetch(2 * (par + 2 * ia));
      a1 = fetch(2 * (par + 2 * ia) + 1);
      pair(b0, b1, hb);
      if (!ha) break;
    }
  }
}

// End of a tile: the two partial sums of each block meet in LDS (one wave
// of a pair stores, the other adds) and the tile goes to HBM once,
// coalesced along z (global atomics for split tiles).  Starts with the
// barrier after which the batch buffers become the staging area.
template <int NDIM, bool CPLX>
__device__ __forceinline__ void mf_flush(const mf32x4 (&acc)[16], const MfLds<NDIM>& S,
                                         const MfLane& ln, const int* origin,
                                         const InterpGeom& G, float* __restrict__ out,
                                         int64_t c0, int nc, bool shared) {
  using T = typename Ch<CPLX>::T;
  constexpr int CH = CPLX ? MF_NPL / 2 : MF_NPL;
  float* stage = S.stage;
  __syncthreads();
  // Lane holds D rows 4 k4 + i of plane n16, i.e. cells (yy, zz) = (k4, i)
  // of each of its blocks:
  //   3-D block (u, v): x = u, y = 4 (v / 2) + k4, z = 4 (v % 2) + i
  //   2-D block (u, v): y = 4 v + k4, z = 4 u + i
  // stage cell index: 3-D 64 x + 8 y + z, 2-D 32 y + z
  auto stage_at = [&](int j, int q) {
    const int u = 2 * j + ln.xpar;
    int cell;
    if constexpr (NDIM == 3) cell = u * 64 + (4 * (q >> 1) + ln.k4) * 8 + 4 * (q & 1);
    else cell = (4 * q + ln.k4) * 32 + 4 * u;
    return reinterpret_cast<float4*>(stage + ln.n16 * MF_PSTR + cell);
  };
  if (ln.par == 0) {
#pragma unroll
    for (int b = 0; b < 16; ++b)
      *stage_at(b >> 2, b & 3) = make_float4(acc[b][0], acc[b][1], acc[b][2], acc[b][3]);
  }
  __syncthreads();
  if (ln.par == 1) {
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      float4* sp = stage_at(b >> 2, b & 3);
      float4 o4 = *sp;
      o4.x += acc[b][0]; o4.y += acc[b][1]; o4.z += acc[b][2]; o4.w += acc[b][3];
      *sp = o4;
    }
  }
  __syncthreads();
  T* __restrict__ o = reinterpret_cast<T*>(out);
  for (int cell = ln.tid; cell < 512; cell += MF_THREADS) {
    int lx, ly, lz;
    if constexpr (NDIM == 3) {
      lx = cell >> 6; ly = (cell >> 3) & 7; lz = cell & 7;
    } else {
      lx = 0; ly = cell >> 5; lz = cell & 31;
    }
    const int64_t gx = origin[0] + lx, gy = origin[1] + ly, gz = origin[2] + lz;
    if (gx >= G.n[0] || gy >= G.n[1] || gz >= G.n[2]) continue;
    const int64_t g0 = (gx * G.n[1] + gy) * G.n[2] + gz;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      if (c >= nc) break;
      const int64_t gi = (c0 + c) * G.ncells + g0;
      if constexpr (CPLX) {
        const float a0 = stage[(2 * c) * MF_PSTR + cell];
        const float a1 = stage[(2 * c + 1) * MF_PSTR + cell];
        if (!shared) {
          o[gi] = make_float2(a0, a1);
        } else {
          unsafeAtomicAdd(out + 2 * gi, a0);
          unsafeAtomicAdd(out + 2 * gi + 1, a1);
        }
      } else {
        const float a0 = stage[c * MF_PSTR + cell];
        if (!shared) o[gi] = a0;
        else unsafeAtomicAdd(out + gi, a0);
      }
    }
  }
}

__device__ __forceinline__ void tile_origin(const BinGeom& B, int tile, int* origin) {
  int r = tile;
  for (int d = 2; d >= 0; --d) {
    origin[d] = (r % B.ntiles[d]) * B.tile[d];
    r /= B.ntiles[d];
  }
}

// One problem.  REC: the points' values come from the point-major records
// (rec, caller order) and their coordinates from the caller's locs, instead
// of the coil-major input and slocs.  Persistent: each workgroup walks
// (work item, channel chunk) units and loads the first batch of its next
// unit while it finishes the current one.
template <int NDIM, bool CPLX, bool REC>
__global__ __launch_bounds__(MF_THREADS) __attribute__((amdgpu_waves_per_eu(3, 4))) void grid_mfma_kernel(
    const float* __restrict__ vals, float* __restrict__ out, int64_t nbatch,
    InterpGeom G, BinGeom B, const int* __restrict__ perm,
    const float* __restrict__ slocs, const int* __restrict__ tlist,
    const int4* __restrict__ items, const int* __restrict__ nitems_ptr,
    const float* __restrict__ locs, const float4* __restrict__ rec) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  constexpr int MB = MfShape<NDIM>::MB;
  constexpr int CH = CPLX ? MF_NPL / 2 : MF_NPL;  // channels per chunk
  static_assert(MB <= MF_THREADS && MB / 16 <= 64, "one point per thread; a lane per pair");
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const MfLds<NDIM> S(lds_raw);
  const MfLane ln;
  const int tid = ln.tid;
  const T* __restrict__ v = reinterpret_cast<const T*>(vals);
  const int nitems = *nitems_ptr;

  // raw inputs of this thread's point of a batch, loaded one batch ahead
  // (across channel chunks and work items too)
  float pt[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) pt[d] = 0.f;
  T pv[CH];
  auto load = [&](int start, int64_t c0, int b0, int nb) {
    if (tid < nb) {
      const int q = tlist[start + b0 + tid];
      if constexpr (REC) {
        const int64_t p = perm[q];
        const float4* r = rec + ((c0 / CH) * B.npts + p) * MF_REC4;
#pragma unroll
        for (int d = 0; d < NDIM; ++d) pt[d] = locs[p * NDIM + d];
        float f[MF_NPL];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 x = r[i];
          f[4 * i] = x.x; f[4 * i + 1] = x.y; f[4 * i + 2] = x.z; f[4 * i + 3] = x.w;
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
          if constexpr (CPLX) pv[c] = make_float2(f[2 * c], f[2 * c + 1]);
          else pv[c] = f[c];
        }
      } else {
#pragma unroll
        for (int d = 0; d < NDIM; ++d) pt[d] = slocs[(int64_t)q * NDIM + d];
        const int p = perm[q];
#pragma unroll
        for (int c = 0; c < CH; ++c)
          pv[c] = c0 + c < nbatch ? v[(c0 + c) * B.npts + p] : C::zero();
      }
    }
  };
  int64_t it = xcd_remap(blockIdx.x, gridDim.x);
  if (it >= nitems) return;
  int4 item = items[it];
  int64_t c0 = 0;
  load(item.y, 0, 0, item.z < MB ? item.z : MB);
  for (;;) {
    const int start = item.y, count = item.z;
    const int nc = (int)(nbatch - c0 < CH ? nbatch - c0 : CH);
    // the next (item, channel chunk) of this workgroup
    int64_t it_n = it, c0_n = c0 + CH;
    int4 item_n = item;
    if (c0_n >= nbatch) {
      c0_n = 0;
      it_n = it + gridDim.x;
      if (it_n < nitems) item_n = items[it_n];
    }
    const bool has_next = it_n < nitems;
    int origin[3];
    tile_origin(B, item.x, origin);
    mf32x4 acc[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) acc[b] = (mf32x4){0.f, 0.f, 0.f, 0.f};
    bool ahead = false;  // the next unit's first batch is loaded
    for (int b0 = 0; b0 < count; b0 += MB) {
      const int nb = count - b0 < MB ? count - b0 : MB;
      __syncthreads();  // previous batch's MFMA reads (or the write-out) are done
      mf_phase1<NDIM, CPLX>(S, ln, nb, pt, pv, origin, G);
      if (b0 + MB < count) {
        load(start, c0, b0 + MB, count - b0 - MB < MB ? count - b0 - MB : MB);
      } else if (has_next) {
        load(item_n.y, c0_n, 0, item_n.z < MB ? item_n.z : MB);
        ahead = true;
      }
      __syncthreads();
      mf_phase2<NDIM>(acc, S, ln, nb);
    }
    if (!ahead && has_next) load(item_n.y, c0_n, 0, item_n.z < MB ? item_n.z : MB);
    mf_flush<NDIM, CPLX>(acc, S, ln, origin, G, out, c0, nc, item.w != 0);
    if (!has_next) break;
    it = it_n;
    item = item_n;
    c0 = c0_n;
  }
}

// Bank cost of one gridding RMW instruction: lane l touches the cell at
// offset X*k0 + Y*k1 + k2 from the point's first tap (3-D lanes: 4x4x4 taps,
// 2-D: 8x8).  Sum of the worst bank multiplicities, for LDS serving a whole
// wave over 64 banks and for LDS serving half waves over 32 banks.
static int grid_bank_cost(int ndim, int X, int Y) {
  int off[64];
  for (int l = 0; l < 64; ++l)
    off[l] = ndim >= 3 ? X * (l >> 4) + Y * ((l >> 2) & 3) + (l & 3)
                       : Y * (l >> 3) + (l & 7);
  int cost = 0;
  for (int banks = 32; banks <= 64; banks *= 2) {
    for (int g = 0; g < 64; g += banks) {
      int cnt[64] = {0}, worst = 0;
      for (int l = g; l < g + banks; ++l) {
        const int b = ++cnt[off[l] % banks];
        worst = b > worst ? b : worst;
      }
      cost += worst;
    }
  }
  return cost;
}

// Row (Y) and plane (X) strides of the gridding accumulator: the padding
// that minimises bank cost x cells (LDS time per instruction against LDS
// space, i.e. channels per workgroup).  Memoised per tile shape.
static void pick_grid_strides(int ndim, const int tile[3], int* gys, int* gxs) {
  struct Memo { int ndim, t0, t1, t2, y, x; };
  static thread_local Memo memo[8];
  static thread_local int nmemo = 0;
  for (int i = 0; i < nmemo; ++i)
    if (memo[i].ndim == ndim && memo[i].t0 == tile[0] && memo[i].t1 == tile[1] &&
        memo[i].t2 == tile[2]) {
      *gys = memo[i].y;
      *gxs = memo[i].x;
      return;
    }
  int64_t best = INT64_MAX;
  *gys = tile[2];
  *gxs = tile[1] * tile[2];
  for (int Y = tile[2]; Y <= tile[2] + 16; ++Y) {
    const int xmax = ndim >= 3 ? tile[1] * Y + 32 : tile[1] * Y;
    for (int X = tile[1] * Y; X <= xmax; ++X) {
      const int64_t score = (int64_t)grid_bank_cost(ndim, X, Y) * tile[0] * X;
      if (score < best) {
        best = score;
        *gys = Y;
        *gxs = X;
      }
    }
  }
  memo[nmemo % 8] = Memo{ndim, tile[0], tile[1], tile[2], *gys, *gxs};
  ++nmemo;
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
  if (adjoint) pick_grid_strides(L.ndim, B.tile, &B.gys, &B.gxs);
  B.npts = L.npts;
  B.nbins = L.u_nbins;
  return B;
}

static size_t ungrid_lds_bytes(const BinGeom& B, int cc, int comp_bytes) {
  return (size_t)B.wcells * cc * comp_bytes;
}

static size_t grid_lds_bytes(const BinGeom& B, int planes, int nt) {
  return (size_t)grid_lds_layout(B, planes, nt).bytes;
}

// The plan resolves the accumulator (matrix cores for 3-D 8x8x8 and 2-D
// 16x32 tiles unless asked otherwise).
static bool use_grid_mfma(const InterpGeom& G, const tnl_plan_layout& L) {
  return G.ndim >= 2 && L.grid_accum == TNL_GRID_ACCUM_MFMA;
}

bool binned_supported(const InterpGeom& G, const tnl_plan_layout& L,
                      bool adjoint) {
  if (G.ndim < 2 || G.norm != 1) return false;
  for (int d = 0; d < 3; ++d)
    if (G.ntap[d] > 8 || G.ntap[d] != L.ntaps[d]) return false;
  if (adjoint && use_grid_mfma(G, L)) {
    for (int d = 0; d < 3; ++d)
      if (G.n[d] < G.ntap[d]) return false;  // a footprint would self-wrap
    return true;
  }
  if (adjoint) {
    const BinGeom B = make_bin_geom(L, true);
    for (int d = 3 - G.ndim; d < 3; ++d) {
      // gridding lanes cover 5 taps per axis in 3-D, 8 in 2-D
      if (G.ntap[d] > (G.ndim == 3 ? 5 : 8)) return false;
      if (G.n[d] < G.ntap[d]) return false;  // a footprint would self-wrap
      if (B.tile[d] >= (1 << 15)) return false;  // lo is packed in 16 bits
    }
    return (int64_t)grid_lds_bytes(B, 1, G.ndim == 3 ? 5 : 8) <= L.lds_budget;
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

// Workgroups of a persistent kernel that fit on the device at once.
static int resident_workgroups(const void* kern, int threads, size_t lds) {
  int per_cu = 0, dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
  }
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) !=
          hipSuccess || per_cu < 1)
    per_cu = 1;
  return per_cu * ncu;
}

template <int NDIM, bool CPLX>
static int launch_grid_mfma(const float* vals, const float* locs, float* out, int64_t nbatch,
                            const InterpGeom& G, const tnl_plan_layout& L,
                            const char* ws, float* scratch, hipStream_t s) {
  const BinGeom B = make_bin_geom(L, true);
  const int* nitems = (const int*)(ws + L.off_itemoff) + L.g_ntiles_total;
  const int4* items = (const int4*)(ws + L.off_items);
  if (int rc = zero_split_tiles<CPLX>(out, nbatch, G, B, L, ws, s)) return rc;
  float4* rec = reinterpret_cast<float4*>(scratch);
  if (rec) {
    const int64_t ch = CPLX ? MF_NPL / 2 : MF_NPL;
    const int64_t nb = (L.npts + PACK_PTS - 1) / PACK_PTS * ((nbatch + ch - 1) / ch);
    pack_points_kernel<CPLX><<<(unsigned)(nb < 16384 ? nb : 16384), 256, 0, s>>>(
        vals, L.npts, nbatch, rec);
    TNL_LAUNCH_CHECK("pack_points");
  }
  if (NDIM == 3 && rec && grid_wave_ok(L)) {
    const int64_t ch = CPLX ? MF_NPL / 2 : MF_NPL;
    float* wrec = scratch + (nbatch + ch - 1) / ch * L.npts * MF_REC4 * 4;
    return grid_wave_launch(vals, locs, out, nbatch, G, L, ws, scratch, wrec, CPLX ? 1 : 0, s);
  }
  const auto kern = rec ? grid_mfma_kernel<NDIM, CPLX, true> : grid_mfma_kernel<NDIM, CPLX, false>;
  constexpr size_t lds = mf_lds_bytes<NDIM>();
  if (int rc = allow_lds((const void*)kern, lds)) return rc;
  int64_t nwg = resident_workgroups((const void*)kern, MF_THREADS, lds);
  nwg = (nwg + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  const int64_t need = (L.item_capacity + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  if (nwg > need) nwg = need;
  kern<<<(unsigned)nwg, MF_THREADS, lds, s>>>(
      vals, out, nbatch, G, B, (const int*)(ws + L.off_perm),
      (const float*)(ws + L.off_slocs), (const int*)(ws + L.off_tlist), items, nitems, locs,
      rec);
  TNL_LAUNCH_CHECK("grid_mfma");
  return TNL_OK;
}

int64_t grid_scratch_bytes(const tnl_plan_layout& L, int64_t nbatch, int is_complex) {
  if (L.ndim < 2 || L.grid_accum != TNL_GRID_ACCUM_MFMA || nbatch <= 0 || L.npts <= 0)
    return 0;
  const int64_t ch = is_complex ? MF_NPL / 2 : MF_NPL;
  const int64_t recs = (nbatch + ch - 1) / ch * L.npts * MF_REC4 * (int64_t)sizeof(float4);
  return recs + (grid_wave_ok(L) ? grid_wave_wrec_bytes(L) : 0);
}

template <int NDIM, bool CPLX, int CC, int NT>
static int launch_binned(bool adjoint, const float* vals, float* out,
                         int64_t nbatch, const InterpGeom& G,
                         const tnl_plan_layout& L, const char* ws, float* pmb,
                         hipStream_t s) {
  const BinGeom B = make_bin_geom(L, adjoint);
  const int comp_bytes = CPLX ? 8 : 4;
  const int* perm = (const int*)(ws + L.off_perm);
  const float* slocs = (const float*)(ws + L.off_slocs);
  if (!adjoint) {
    const int64_t nwg = (L.sub_capacity + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const size_t lds = ungrid_lds_bytes(B, CC, comp_bytes);
    auto kern = ungrid_binned_kernel<NDIM, CPLX, CC, NT, false>;
    if constexpr (NDIM == 3 && CPLX && CC == 2 && NT == 5) {
      bool ilv = true;
      for (int d = 0; d < 3; ++d) ilv = ilv && B.win[d] == ILV_W && G.n[d] >= ILV_W;
      // z pairs of cells: even n[2], 16-byte aligned channel planes
      ilv = ilv && G.n[2] % 2 == 0 && (reinterpret_cast<uintptr_t>(vals) & 15) == 0;
      if (ilv) kern = ungrid_binned_kernel<3, true, 2, 5, true>;
    }
    if (int rc = allow_lds((const void*)kern, lds)) return rc;
    kern<<<dim3((unsigned)nwg, 1), BIN_THREADS, lds, s>>>(
        vals, pmb ? pmb : out, nbatch, G, B, perm, slocs, (const int4*)(ws + L.off_sub),
        (const int*)(ws + L.off_suboff) + L.u_nbins + 1, pmb != nullptr);
    TNL_LAUNCH_CHECK("ungrid_binned");
    if (pmb) {
      int64_t nwu = (B.npts + UNPK_P - 1) / UNPK_P;
      if (nwu > 16384) nwu = 16384;
      unpack_points_kernel<CPLX><<<(unsigned)nwu, 256, 0, s>>>(pmb, out, nbatch, B.npts);
      TNL_LAUNCH_CHECK("unpack_points");
    }
  } else if constexpr (NDIM == 3 && NT > 5) {
    return fail(TNL_EINVAL, "3-D gridding covers at most 5 taps per axis");
  } else {
    const int* nitems = (const int*)(ws + L.off_itemoff) + L.g_ntiles_total;
    const int4* items = (const int4*)(ws + L.off_items);
    if (int rc = zero_split_tiles<CPLX>(out, nbatch, G, B, L, ws, s)) return rc;
    constexpr int NP = CC * (CPLX ? 2 : 1);
    constexpr int NW = NP >= 4 ? 4 : NP;
    const auto kern = grid_binned_kernel<NDIM, CPLX, CC, NT, NW>;
    const size_t lds = grid_lds_bytes(B, NP, NT);
    if (int rc = allow_lds((const void*)kern, lds)) return rc;
    // persistent: as many workgroups as can be resident walk the items
    int64_t nwg = resident_workgroups((const void*)kern, GW * NW, lds);
    nwg = (nwg + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const int64_t need = (L.item_capacity + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    if (nwg > need) nwg = need;
    kern<<<(unsigned)nwg, GW * NW, lds, s>>>(vals, out, nbatch, G, B, perm, slocs,
                                             (const int*)(ws + L.off_tlist), items, nitems);
    TNL_LAUNCH_CHECK("grid_binned");
  }
  return TNL_OK;
}

template <int NDIM, bool CPLX, int NT>
static int dispatch_cc(bool adjoint, const float* vals, float* out,
                       int64_t nbatch, const InterpGeom& G,
                       const tnl_plan_layout& L, const char* ws, float* pmb,
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
    case 8: return launch_binned<NDIM, CPLX, 8, NT>(adjoint, vals, out, nbatch, G, L, ws, pmb, s);
    case 4: return launch_binned<NDIM, CPLX, 4, NT>(adjoint, vals, out, nbatch, G, L, ws, pmb, s);
    case 2: return launch_binned<NDIM, CPLX, 2, NT>(adjoint, vals, out, nbatch, G, L, ws, pmb, s);
    default: return launch_binned<NDIM, CPLX, 1, NT>(adjoint, vals, out, nbatch, G, L, ws, pmb, s);
  }
}

template <int NDIM, bool CPLX>
static int dispatch_nt(bool adjoint, const float* vals, const float* locs, float* out,
                       int64_t nbatch, const InterpGeom& G,
                       const tnl_plan_layout& L, const char* ws, float* scratch,
                       hipStream_t s) {
  int nt = G.ntap[0];
  for (int d = 1; d < 3; ++d) nt = G.ntap[d] > nt ? G.ntap[d] : nt;
  if (adjoint && use_grid_mfma(G, L))
    return launch_grid_mfma<NDIM, CPLX>(vals, locs, out, nbatch, G, L, ws, scratch, s);
  // interpolation with a scratch: point-major into it, then one transpose
  float* pmb = adjoint ? nullptr : scratch;
  if (nt <= 5) return dispatch_cc<NDIM, CPLX, 5>(adjoint, vals, out, nbatch, G, L, ws, pmb, s);
  return dispatch_cc<NDIM, CPLX, 8>(adjoint, vals, out, nbatch, G, L, ws, pmb, s);
}

int interp_binned(bool adjoint, const float* vals, const float* locs, float* out,
                  int64_t nbatch, const InterpGeom& G, int is_complex,
                  const tnl_plan_layout& L, const void* plan, float* scratch,
                  hipStream_t s) {
  const char* ws = (const char*)plan;
  if (G.ndim == 3)
    return is_complex
               ? dispatch_nt<3, true>(adjoint, vals, locs, out, nbatch, G, L, ws, scratch, s)
               : dispatch_nt<3, false>(adjoint, vals, locs, out, nbatch, G, L, ws, scratch, s);
  return is_complex
             ? dispatch_nt<2, true>(adjoint, vals, locs, out, nbatch, G, L, ws, scratch, s)
             : dispatch_nt<2, false>(adjoint, vals, locs, out, nbatch, G, L, ws, scratch, s);
}

// ---------------------------------------------------------------- batches
// Device table entry of one problem of a batch.  Global work indices are
// the concatenation of the problems' subproblem (ungrid) and work-item
// (grid) capacities; unused capacity is skipped.
struct BatchDesc {
  InterpGeom G;
  BinGeom Bu, Bg;
  int64_t nbatch;
  int64_t grid_off, pts_off;  // float offsets into the base buffers
  int64_t sub_begin, item_begin;
  const int* perm;
  const float* slocs;
  const int4* sub;
  const int* nsub;
  const int* tlist;
  const int4* items;
  const int* nitems;
  const int* gstarts;
};

// Last problem whose range starts at or before g (g is wave-uniform).
__device__ __forceinline__ int find_problem(const BatchDesc* __restrict__ D,
                                            int nprob, int64_t g, bool items) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int64_t b = items ? D[mid].item_begin : D[mid].sub_begin;
    if (b <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <int NDIM, bool CPLX, int CC, int NT>
__global__ __launch_bounds__(BIN_THREADS) void ungrid_batched_kernel(
    const BatchDesc* __restrict__ D, int nprob, int64_t total,
    const float* __restrict__ gbase, float* __restrict__ pbase) {
  const int64_t s = xcd_remap(blockIdx.x, gridDim.x);
  if (s >= total) return;
  const BatchDesc& d = D[find_problem(D, nprob, s, false)];
  const int64_t sl = s - d.sub_begin;
  if (sl >= *d.nsub) return;
  ungrid_sub<NDIM, CPLX, CC, NT>(gbase + d.grid_off, pbase + d.pts_off, d.nbatch,
                                 d.G, d.Bu, d.perm, d.slocs, d.sub[sl]);
}

template <bool CPLX>
__global__ void grid_zero_shared_batched_kernel(const BatchDesc* __restrict__ D,
                                                int nprob, int64_t total,
                                                float* __restrict__ gbase) {
  for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
    const BatchDesc& d = D[find_problem(D, nprob, g, true)];
    const int64_t it = g - d.item_begin;
    if (it >= *d.nitems) continue;
    zero_shared_item<CPLX>(gbase + d.grid_off, d.nbatch, d.G, d.Bg, d.items[it],
                           d.gstarts);
  }
}

template <int NDIM, bool CPLX, int CC, int NT, int NW>
__global__ __launch_bounds__(GW * NW) void grid_batched_kernel(
    const BatchDesc* __restrict__ D, int nprob, int64_t total,
    const float* __restrict__ pbase, float* __restrict__ gbase) {
  const GridLanes ln = make_grid_lanes<NDIM>(threadIdx.x & (GW - 1));
  for (int64_t g = xcd_remap(blockIdx.x, gridDim.x); g < total; g += gridDim.x) {
    const BatchDesc& d = D[find_problem(D, nprob, g, true)];
    const int64_t it = g - d.item_begin;
    if (it >= *d.nitems) continue;
    grid_item<NDIM, CPLX, CC, NT, NW>(ln, pbase + d.pts_off, gbase + d.grid_off,
                                      d.nbatch, d.G, d.Bg, d.perm, d.slocs,
                                      d.tlist, d.items[it]);
  }
}


int batch_prepare(const InterpGeom* G, const tnl_batch_problem* P, int nprob,
                  int is_complex, tnl_batch_layout* L) {
  const int comp_bytes = is_complex ? 8 : 4;
  const int comp = is_complex ? 2 : 1;
  int nt = 1;
  int64_t budget = INT64_MAX;
  L->total_sub = L->total_items = 0;
  for (int i = 0; i < nprob; ++i) {
    const tnl_plan_layout& pl = *P[i].layout;
    // tiles may differ per problem (plans clamp them to each grid); the
    // launches take the largest LDS need
    if (pl.ndim != P[0].layout->ndim || pl.pad_mode != P[0].layout->pad_mode)
      return fail(TNL_EINVAL, "every plan of a batch must have the same ndim and pad_mode");
    if (!binned_supported(G[i], pl, false) || !binned_supported(G[i], pl, true))
      return fail(TNL_EINVAL,
                  "batched problems must fit the plan-driven kernels (2-D/3-D, norm 1, "
                  "taps per axis within the limits)");
    for (int d = 0; d < 3; ++d) nt = G[i].ntap[d] > nt ? G[i].ntap[d] : nt;
    budget = pl.lds_budget < budget ? pl.lds_budget : budget;
    L->total_sub += pl.sub_capacity;
    L->total_items += pl.item_capacity;
  }
  L->nt = nt <= 5 ? 5 : 8;
  L->grid_mfma = 1;
  for (int i = 0; i < nprob; ++i)
    if (P[i].layout->grid_accum != TNL_GRID_ACCUM_MFMA) L->grid_mfma = 0;
  // the matrix-core gridding takes any tap count up to 8; the LDS
  // accumulator's lanes cover 5 per axis in 3-D
  if (L->ndim == 3 && L->nt > 5 && !L->grid_mfma)
    return fail(TNL_EINVAL,
                "3-D batches wider than 5 taps per axis need matrix-core gridding plans");
  L->reserved = 0;
  // channels per workgroup: the largest power of two (<= 8) whose LDS fits
  // every problem under the smallest budget
  auto max_u = [&](int cc) {
    int64_t m = 0;
    for (int i = 0; i < nprob; ++i) {
      const int64_t b = (int64_t)ungrid_lds_bytes(make_bin_geom(*P[i].layout, false), cc, comp_bytes);
      m = b > m ? b : m;
    }
    return m;
  };
  auto max_g = [&](int cc) {
    int64_t m = 0;
    for (int i = 0; i < nprob; ++i) {
      const int64_t b = (int64_t)grid_lds_bytes(make_bin_geom(*P[i].layout, true), cc * comp, L->nt);
      m = b > m ? b : m;
    }
    return m;
  };
  int64_t maxc = 1;
  for (int i = 0; i < nprob; ++i) maxc = P[i].nbatch > maxc ? P[i].nbatch : maxc;
  int ucc = 1, gcc = 1;
  while (ucc < 8 && ucc < maxc && max_u(2 * ucc) <= budget) ucc *= 2;
  while (gcc < 8 && gcc < maxc && max_g(2 * gcc) <= budget) gcc *= 2;
  // one channel per workgroup is always taken (as for a single problem);
  // the hard limit is the CU's LDS
  if (max_u(ucc) > 160 * 1024 || max_g(gcc) > 160 * 1024)
    return fail(TNL_EINVAL, "batch needs more LDS per workgroup than a CU has");
  L->ungrid_cc = ucc;
  L->grid_cc = gcc;
  L->ungrid_lds = max_u(ucc);
  L->grid_lds = max_g(gcc);
  L->ungrid_chunks = (int32_t)((maxc + ucc - 1) / ucc);
  L->bytes = (int64_t)nprob * (int64_t)sizeof(BatchDesc);
  (void)comp;
  return TNL_OK;
}

int batch_build(const InterpGeom* G, const tnl_batch_problem* P,
                const tnl_batch_layout& L, void* ws, hipStream_t s) {
  const int comp = L.is_complex ? 2 : 1;
  BatchDesc* D = (BatchDesc*)std::calloc((size_t)L.nprob, sizeof(BatchDesc));
  if (!D) return fail(TNL_EINVAL, "out of host memory");
  int64_t sub_begin = 0, item_begin = 0;
  for (int i = 0; i < L.nprob; ++i) {
    const tnl_plan_layout& pl = *P[i].layout;
    const char* w = (const char*)P[i].plan;
    BatchDesc& d = D[i];
    d.G = G[i];
    d.Bu = make_bin_geom(pl, false);
    d.Bg = make_bin_geom(pl, true);
    d.nbatch = P[i].nbatch;
    d.grid_off = P[i].grid_offset * comp;
    d.pts_off = P[i].pts_offset * comp;
    d.sub_begin = sub_begin;
    d.item_begin = item_begin;
    sub_begin += pl.sub_capacity;
    item_begin += pl.item_capacity;
    d.perm = (const int*)(w + pl.off_perm);
    d.slocs = (const float*)(w + pl.off_slocs);
    d.sub = (const int4*)(w + pl.off_sub);
    d.nsub = (const int*)(w + pl.off_suboff) + pl.u_nbins + 1;
    d.tlist = (const int*)(w + pl.off_tlist);
    d.items = (const int4*)(w + pl.off_items);
    d.nitems = (const int*)(w + pl.off_itemoff) + pl.g_ntiles_total;
    d.gstarts = (const int*)(w + pl.off_gstarts);
  }
  int rc = check_hip(hipMemcpyAsync(ws, D, (size_t)L.bytes, hipMemcpyHostToDevice, s),
                     "batch table copy");
  if (!rc) rc = check_hip(hipStreamSynchronize(s), "batch table copy");
  std::free(D);
  return rc;
}

template <int NDIM, bool CPLX, int CC, int NT>
static int launch_ungrid_batched(const tnl_batch_layout& L, const BatchDesc* D,
                                 const float* in, float* out, hipStream_t s) {
  const auto kern = ungrid_batched_kernel<NDIM, CPLX, CC, NT>;
  if (int rc = allow_lds((const void*)kern, (size_t)L.ungrid_lds)) return rc;
  const int64_t nwg = (L.total_sub + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  kern<<<dim3((unsigned)nwg, 1), BIN_THREADS, (size_t)L.ungrid_lds, s>>>(D, L.nprob,
                                                                        L.total_sub, in, out);
  TNL_LAUNCH_CHECK("ungrid_batched");
  return TNL_OK;
}

// Matrix-core gridding of a batch: one persistent launch walks the work
// items of every problem (global item index -> problem by binary search,
// unused capacity skipped), each item with all its channel chunks, through
// the phases of grid_mfma_kernel.  Values are read from each problem's
// coil-major input (no point-major records).
template <int NDIM, bool CPLX>
__global__ __launch_bounds__(MF_THREADS) __attribute__((amdgpu_waves_per_eu(3, 4))) void grid_mfma_batched_kernel(
    const BatchDesc* __restrict__ D, int nprob, int64_t total,
    const float* __restrict__ pbase, float* __restrict__ gbase) {
  using C = Ch<CPLX>;
  using T = typename C::T;
  constexpr int MB = MfShape<NDIM>::MB;
  constexpr int CH = CPLX ? MF_NPL / 2 : MF_NPL;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const MfLds<NDIM> S(lds_raw);
  const MfLane ln;
  float pt[NDIM];
#pragma unroll
  for (int d = 0; d < NDIM; ++d) pt[d] = 0.f;
  T pv[CH];
  for (int64_t g = xcd_remap(blockIdx.x, gridDim.x); g < total; g += gridDim.x) {
    const BatchDesc& d = D[find_problem(D, nprob, g, true)];
    const int64_t il = g - d.item_begin;
    if (il >= *d.nitems) continue;
    const int4 item = d.items[il];
    const int start = item.y, count = item.z;
    const T* __restrict__ v = reinterpret_cast<const T*>(pbase + d.pts_off);
    float* out = gbase + d.grid_off;
    const int64_t nbatch = d.nbatch;
    int origin[3];
    tile_origin(d.Bg, item.x, origin);
    auto load = [&](int64_t c0, int b0, int nb) {
      if (ln.tid < nb) {
        const int q = d.tlist[start + b0 + ln.tid];
#pragma unroll
        for (int k = 0; k < NDIM; ++k) pt[k] = d.slocs[(int64_t)q * NDIM + k];
        const int p = d.perm[q];
#pragma unroll
        for (int c = 0; c < CH; ++c)
          pv[c] = c0 + c < nbatch ? v[(c0 + c) * d.Bg.npts + p] : C::zero();
      }
    };
    for (int64_t c0 = 0; c0 < nbatch; c0 += CH) {
      const int nc = (int)(nbatch - c0 < CH ? nbatch - c0 : CH);
      mf32x4 acc[16];
#pragma unroll
      for (int b = 0; b < 16; ++b) acc[b] = (mf32x4){0.f, 0.f, 0.f, 0.f};
      load(c0, 0, count < MB ? count : MB);
      for (int b0 = 0; b0 < count; b0 += MB) {
        const int nb = count - b0 < MB ? count - b0 : MB;
        __syncthreads();  // previous batch's MFMA reads (or the write-out) are done
        mf_phase1<NDIM, CPLX>(S, ln, nb, pt, pv, origin, d.G);
        if (b0 + MB < count) load(c0, b0 + MB, count - b0 - MB < MB ? count - b0 - MB : MB);
        __syncthreads();
        mf_phase2<NDIM>(acc, S, ln, nb);
      }
      mf_flush<NDIM, CPLX>(acc, S, ln, origin, d.G, out, c0, nc, item.w != 0);
    }
  }
}

template <int NDIM, bool CPLX>
static int launch_grid_mfma_batched(const tnl_batch_layout& L, const BatchDesc* D,
                                    const float* in, float* out, hipStream_t s) {
  grid_zero_shared_batched_kernel<CPLX><<<1024, 256, 0, s>>>(D, L.nprob, L.total_items, out);
  TNL_LAUNCH_CHECK("grid_zero_shared_batched");
  const auto kern = grid_mfma_batched_kernel<NDIM, CPLX>;
  constexpr size_t lds = mf_lds_bytes<NDIM>();
  if (int rc = allow_lds((const void*)kern, lds)) return rc;
  int64_t nwg = resident_workgroups((const void*)kern, MF_THREADS, lds);
  nwg = (nwg + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  const int64_t need = (L.total_items + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
  if (nwg > need) nwg = need;
  kern<<<(unsigned)nwg, MF_THREADS, lds, s>>>(D, L.nprob, L.total_items, in, out);
  TNL_LAUNCH_CHECK("grid_mfma_batched");
  return TNL_OK;
}

template <int NDIM, bool CPLX, int CC, int NT>
static int launch_grid_batched(const tnl_batch_layout& L, const BatchDesc* D,
                               const float* in, float* out, hipStream_t s) {
  constexpr int NP = CC * (CPLX ? 2 : 1);
  constexpr int NW = NP >= 4 ? 4 : NP;
  if constexpr (NDIM == 3 && NT > 5) {
    return fail(TNL_EINVAL, "3-D gridding covers at most 5 taps per axis");
  } else {
    grid_zero_shared_batched_kernel<CPLX><<<1024, 256, 0, s>>>(D, L.nprob, L.total_items,
                                                                out);
    TNL_LAUNCH_CHECK("grid_zero_shared_batched");
    const auto kern = grid_batched_kernel<NDIM, CPLX, CC, NT, NW>;
    if (int rc = allow_lds((const void*)kern, (size_t)L.grid_lds)) return rc;
    int64_t nwg = resident_workgroups((const void*)kern, GW * NW, (size_t)L.grid_lds);
    nwg = (nwg + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    const int64_t need = (L.total_items + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    if (nwg > need) nwg = need;
    kern<<<(unsigned)nwg, GW * NW, (size_t)L.grid_lds, s>>>(D, L.nprob, L.total_items, in,
                                                            out);
    TNL_LAUNCH_CHECK("grid_batched");
    return TNL_OK;
  }
}

template <int NDIM, bool CPLX, int NT>
static int batch_dispatch_cc(bool adjoint, const tnl_batch_layout& L,
                             const BatchDesc* D, const float* in, float* out,
                             hipStream_t s) {
  const int cc = adjoint ? L.grid_cc : L.ungrid_cc;
#define TNL_BATCH_CASE(C)                                                     \
  case C:                                                                     \
    return adjoint ? launch_grid_batched<NDIM, CPLX, C, NT>(L, D, in, out, s) \
                   : launch_ungrid_batched<NDIM, CPLX, C, NT>(L, D, in, out, s);
  switch (cc) {
    TNL_BATCH_CASE(8)
    TNL_BATCH_CASE(4)
    TNL_BATCH_CASE(2)
    TNL_BATCH_CASE(1)
    default:
      return fail(TNL_EINVAL, "bad channel chunk in batch layout");
  }
#undef TNL_BATCH_CASE
}

int batch_run(bool adjoint, const tnl_batch_layout& L, const void* ws,
              const float* in, float* out, hipStream_t s) {
  const BatchDesc* D = (const BatchDesc*)ws;
  const bool cx = L.is_complex != 0;
  if (adjoint && L.grid_mfma) {
    if (L.ndim == 3)
      return cx ? launch_grid_mfma_batched<3, true>(L, D, in, out, s)
                : launch_grid_mfma_batched<3, false>(L, D, in, out, s);
    return cx ? launch_grid_mfma_batched<2, true>(L, D, in, out, s)
              : launch_grid_mfma_batched<2, false>(L, D, in, out, s);
  }
  if (L.ndim == 3) {
    if (L.nt == 5)
      return cx ? batch_dispatch_cc<3, true, 5>(adjoint, L, D, in, out, s)
                : batch_dispatch_cc<3, false, 5>(adjoint, L, D, in, out, s);
    // wider 3-D batches: interpolation here, gridding on the matrix cores
    // above (the layout refuses them otherwise)
    return cx ? batch_dispatch_cc<3, true, 8>(adjoint, L, D, in, out, s)
              : batch_dispatch_cc<3, false, 8>(adjoint, L, D, in, out, s);
  }
  if (L.nt == 5)
    return cx ? batch_dispatch_cc<2, true, 5>(adjoint, L, D, in, out, s)
              : batch_dispatch_cc<2, false, 5>(adjoint, L, D, in, out, s);
  return cx ? batch_dispatch_cc<2, true, 8>(adjoint, L, D, in, out, s)
            : batch_dispatch_cc<2, false, 8>(adjoint, L, D, in, out, s);
}

}  // namespace tnl
