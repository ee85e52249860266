// 3-D gridding on the matrix cores, one wave per work item ("wave
// gridding").  Used for plans whose gridding tiles are 8x8x8 cells and whose
// work items carry (point, block mask) touch pairs (plan sections `tpm`),
// when the caller passes the scratch of tnl_grid_ws.
//
// Per call, two pre-passes fill the scratch:
//   * point-major channel records (pack_points_kernel, shared with the
//     4-wave kernel): 16 planes (8 complex / 16 real channels) per point and
//     channel chunk, 64 bytes;
//   * window records (window_rec_kernel): per point the weights of its kept
//     taps on each axis (5 slots per axis, zero past the kept taps) and the
//     first kept tap of each axis wrapped into [0, n), 10 bits each; 64
//     bytes.
// A wave owns one (work item, channel chunk) and the whole 8x8x8 x 16-plane
// accumulator of its tile in registers: 32 blocks of 16 cells (one x plane,
// a 4x4 (y, z) patch) x 16 planes, each one v_mfma_f32_16x16x4f32 tile.
// Touches are taken 4 at a time (the MFMA's K): lane (k, n) = (lane / 16,
// lane % 16) belongs to touch k, plane n.  Separability splits the operands:
//   A[cell (y, z) of patch P][k] = wy_k(y) * wz_k(z)    (4 values: 4 patches)
//   B_x[k][n]                    = wx_k(x) * v_k[n]     (8 values: 8 planes)
// so a group costs 12 multiplies for all of its MFMAs.  Touches go in
// batches of 32; both records of a batch's touches come into LDS by direct
// loads.  Each touch's 1-D windows are placed at their tile-local cells
// through a small per-wave LDS table: rows of 16 slots per (touch, axis),
// cleared, the touch's 5 weights written at slot 4 + (first tap - tile
// origin) + k, and slots 4..11 (the tile's 8 cells) read back, so no lane
// selects among registers.  Groups of 4 touches go in pairs: each touch pair
// carries the union of the block masks of its group (written by the plan),
// and the pair's MFMAs cover the union of its two groups.  The records of
// batch b + 1 and the touch pairs of batch b + 2 load while batch b is
// processed.  At the end of the item the
// accumulators go to HBM one x plane at a time through an LDS stage
// (64-byte z runs), or with float atomics for split tiles (their region is
// zeroed by the caller's pre-pass).  No barriers: the waves of a CU never
// share LDS.
#include "tnl_common.hpp"
#include "tnl_internal.hpp"

namespace tnl {

namespace {

typedef float wv32x4 __attribute__((ext_vector_type(4)));

constexpr int WG_NPL = 16;          // planes of a channel chunk
constexpr int WG_REC = 16;          // floats per channel / window record
constexpr int WG_PS = 68;           // LDS stage stride per plane (floats)
constexpr int WG_RUN = 8;           // consecutive work units given to one XCD
constexpr int WG_BG = 8;            // groups of 4 touches per prefetch batch
constexpr int WG_BT = 4 * WG_BG;    // touches per batch (one lane per record)
constexpr int WG_XCD = 8;

// Window record of every point (caller order): weights of the kept taps of
// each axis at slots 5 d + k (zero for k past the kept taps) and, in slot
// 15, the first kept tap of each axis wrapped into [0, n), bits 10 d.  The
// weights are the ones the 4-wave kernel evaluates per touch
// (mf_axis_place): window(min(|(g0 + k) - t| * 2 / w, 1)).
__global__ __launch_bounds__(256) void window_rec_kernel(const float* __restrict__ locs,
                                                         int64_t npts, InterpGeom G,
                                                         float4* __restrict__ wrec) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < npts;
       p += (int64_t)gridDim.x * 256) {
    float w[WG_REC];
    unsigned packed = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float t = locs[p * 3 + d];
      const TapRange r = tap_range(t, G.half[d], G.ntap[d], G.n[d], G.pad);
      int64_t a = r.g0;
      if (G.pad == TNL_PAD_CIRCULAR) {
        if (a < 0) a += G.n[d];
        else if (a >= G.n[d]) a -= G.n[d];
        if (a < 0 || a >= G.n[d]) a = mod_pos(r.g0, G.n[d]);
      }
      if (r.count == 0) a = 0;
      const float g0f = (float)r.g0;
#pragma unroll
      for (int k = 0; k < 5; ++k)
        w[5 * d + k] = k < r.count ? window(G.kernel,
                                            fminf(fabsf((g0f + (float)k) - t) * G.inv_half[d],
                                                  1.0f),
                                            G.beta)
                                   : 0.0f;
      packed |= (unsigned)a << (10 * d);
    }
    w[15] = __uint_as_float(packed);
    float4* o = wrec + p * 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = make_float4(w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]);
  }
}

struct WaveGeom {
  int n[3];
  int ntiles[3];
  int64_t ncells;
  int64_t npts;
  int64_t nbatch;
  int nch;   // channel chunks
  int circ;  // circular padding
};

// One wave per (work item, channel chunk).
template <bool CPLX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 3))) void grid_wave_kernel(
    float* __restrict__ out, WaveGeom W, const int2* __restrict__ tpm,
    const float* __restrict__ wrec, const float* __restrict__ rec,
    const int4* __restrict__ items, const int* __restrict__ nitems_ptr) {
  constexpr int CH = CPLX ? WG_NPL / 2 : WG_NPL;
  // records of two batches of WG_BT touches, as the direct-to-LDS loads
  // leave them: [buffer][16-byte chunk i][lane][4], lane = kind * 32 + touch
  // (kind 0: window record, 1: channel record); the write-out stage reuses
  // them once every load has landed
  // (two separate arrays, so that the compiler's wait analysis can tell
  // the buffer being read from the one being loaded)
  __shared__ __align__(16) float rb0[4][64][4];
  __shared__ __align__(16) float rb1[4][64][4];
  // tile-local windows of the batch's touches: [touch][axis][16 slots],
  // the tile's 8 cells at slots 4..11; slots 0..3 and 12..15 catch the taps
  // that fall outside the tile, so placing a weight needs no test
  __shared__ __align__(16) float tw[WG_BT * 3 * 16];
  __shared__ __align__(16) float stg[WG_NPL * WG_PS];

  const int64_t nunits = (int64_t)(*nitems_ptr) * W.nch;
  // runs of WG_RUN consecutive units on one XCD (the dispatcher hands
  // consecutive workgroups to the XCDs in turn)
  const int64_t bid = blockIdx.x;
  const int64_t rnd = bid / (WG_XCD * WG_RUN), j = bid - rnd * (WG_XCD * WG_RUN);
  const int64_t unit = rnd * (WG_XCD * WG_RUN) + (j % WG_XCD) * WG_RUN + j / WG_XCD;
  if (unit >= nunits) return;
  const int64_t itx = unit / W.nch;
  const int ch = (int)(unit - itx * W.nch);
  const int64_t c0 = (int64_t)ch * CH;
  const int4 item = items[itx];
  const int start = item.y, count = item.z;

  int origin[3];
  {
    int r = item.x;
    for (int d = 2; d >= 0; --d) {
      origin[d] = (r % W.ntiles[d]) * 8;
      r /= W.ntiles[d];
    }
  }

  const int lane = threadIdx.x;
  const int k4 = lane >> 4, n16 = lane & 15;
  const int yy = n16 >> 2, zz = n16 & 3;

  wv32x4 acc[32];
#pragma unroll
  for (int b = 0; b < 32; ++b) acc[b] = (wv32x4){0.f, 0.f, 0.f, 0.f};

  const int nbatches = (count + WG_BT - 1) / WG_BT;
  // Touch pair of this lane's touch (lane % 32) of batch b: the caller
  // index and the union of the block masks of the touch's group of 4 (the
  // plan writes it to all 4 pairs).  Touches past the item's last read its
  // last pair.
  auto pair_of = [&](int b) {
    int i = WG_BT * b + (lane & (WG_BT - 1));
    i = i < count ? i : count - 1;
    return tpm[start + i];
  };
  // Issue the direct-to-LDS loads of batch b's records into buffer buf:
  // lane (kind, touch) copies one 64-byte record in four 16-byte chunks.
  // Touches past the last copy point 0's records (never read).
  auto issue = [&](int b, int2 t, float (*dst)[64][4]) {
    const bool live = WG_BT * b + (lane & (WG_BT - 1)) < count;
    const int64_t p = live ? t.x : 0;
    const float* src = lane < WG_BT ? wrec + p * WG_REC
                                    : rec + ((int64_t)ch * W.npts + p) * WG_REC;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      __builtin_amdgcn_global_load_lds(src + 4 * i, &dst[i][0][0], 16, 0, 0);
  };
  // Place the tile-local windows of the batch in rb (touches 0..left-1)
  // into tw: each (touch, axis) task clears its row, then writes the
  // weights of its kept taps at their cells.  All reads of a lane's tasks
  // are issued before its stores (one LDS round trip, no per-tap waits).
  // per lane and pass j: the task's axis, tile origin and grid size along it
  // (masks, not selects: selects over an array would become a stack load)
  int st_a[2], st_org[2], st_n[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int a = (64 * j + lane) % 3;
    const int m0 = -(a == 0), m1 = -(a == 1), m2 = -(a == 2);
    st_a[j] = a;
    st_org[j] = (origin[0] & m0) | (origin[1] & m1) | (origin[2] & m2);
    st_n[j] = (W.n[0] & m0) | (W.n[1] & m1) | (W.n[2] & m2);
  }
  auto stage_windows = [&](const float (*rb)[64][4], int left) {
    unsigned base[2];
    float wk[2][5];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int task = 64 * j + lane, a = st_a[j];
      const int t = task < 3 * WG_BT ? task / 3 : 0;
      base[j] = __float_as_uint(rb[3][t][3]);
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int wd = 5 * a + k;
        wk[j][k] = rb[wd >> 2][t][wd & 3];
      }
    }
    // every row is cleared, live or not (a row of a dead touch still
    // multiplies a zero value, which must not meet a stale weight); lane l
    // clears 16-byte units l, l + 64, ..., so each store instruction covers
    // 1 KB without bank conflicts (clearing row by row put 16 lanes on the
    // same banks)
#pragma unroll
    for (int i = 0; i < WG_BT * 3 * 16 / 4 / 64; ++i)
      reinterpret_cast<float4*>(tw)[64 * i + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int task = 64 * j + lane, t = task / 3, a = st_a[j];
      if (task < 3 * WG_BT) {
        float* row = tw + (t * 3 + a) * 16;
        if (t < left) {
          int d = (int)((base[j] >> (10 * a)) & 1023u) - st_org[j];
          if (W.circ && d > 7) d -= st_n[j];
          // a touch's first tap lies 4 cells before the tile at most and
          // inside it at the latest; the clamp keeps the stores in the row
          d = d < -4 ? -4 : (d > 7 ? 7 : d);
#pragma unroll
          for (int k = 0; k < 5; ++k) row[4 + d + k] = wk[j][k];
        }
      }
    }
  };
  // Operands of group gg of the staged batch: this lane's touch 4 gg + k4.
  struct Ops {
    float4 x0, x1;           // x window at the tile's 8 planes
    float wy0, wy1, wz0, wz1;  // y window at rows yy, yy + 4; z at zz, zz + 4
    float v;                 // channel value of plane n16
  };
  auto ops_of = [&](const float (*rb)[64][4], int gg, bool live) {
    const int t = 4 * gg + k4;
    const float* r = tw + t * 48 + 4;  // x, y, z rows at r, r + 16, r + 32
    Ops o;
    o.x0 = *reinterpret_cast<const float4*>(r);
    o.x1 = *reinterpret_cast<const float4*>(r + 4);
    o.wy0 = r[16 + yy];
    o.wy1 = r[20 + yy];
    o.wz0 = r[32 + zz];
    o.wz1 = r[36 + zz];
    o.v = live ? rb[n16 >> 2][WG_BT + t][n16 & 3] : 0.f;
    return o;
  };
  // The MFMAs of two groups (8 touches) under one decision per block: A (4
  // patches) from the y/z windows and B (one per live plane) from the x
  // window and the value, for each group; gm is the union of both groups'
  // block masks (a group absent from a block adds zeros there).
  auto mfmas = [&](const Ops& o, const Ops& p, unsigned gm) {
    const float Ao[4] = {o.wy0 * o.wz0, o.wy0 * o.wz1, o.wy1 * o.wz0, o.wy1 * o.wz1};
    const float Ap[4] = {p.wy0 * p.wz0, p.wy0 * p.wz1, p.wy1 * p.wz0, p.wy1 * p.wz1};
    const float xo[8] = {o.x0.x, o.x0.y, o.x0.z, o.x0.w, o.x1.x, o.x1.y, o.x1.z, o.x1.w};
    const float xp[8] = {p.x0.x, p.x0.y, p.x0.z, p.x0.w, p.x1.x, p.x1.y, p.x1.z, p.x1.w};
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const unsigned pm = (gm >> (4 * u)) & 15u;
      if (pm) {
        const float bo = xo[u] * o.v, bp = xp[u] * p.v;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if ((pm >> q) & 1u) {
            acc[4 * u + q] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ao[q], bo, acc[4 * u + q], 0, 0, 0);
            acc[4 * u + q] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ap[q], bp, acc[4 * u + q], 0, 0, 0);
          }
      }
    }
  };

  // Batch b + 1's records land in LDS and batch b + 2's touch pairs load
  // while batch b is processed.  Past the last batch the loads are clamped
  // repeats (never read).
  {
    int2 tcur = pair_of(0);
    int2 tnext = pair_of(1);
    issue(0, tcur, rb0);
    auto half = [&](int b, const float (*rd)[64][4], float (*wr)[64][4]) {
      // batch b's records and batch b + 1's pairs, issued one batch ago
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
      issue(b + 1 < nbatches ? b + 1 : nbatches - 1, tnext, wr);
      const int2 tnn = pair_of(b + 2);
      const int left = count - WG_BT * b;
      const int ng = left <= 0 ? 0 : (left >= WG_BT ? WG_BG : (left + 3) >> 2);
      if (ng > 0) {
        stage_windows(rd, left);
        // groups in pairs; a missing second group has no live touch (its
        // operands are zero) and no mask bits
#pragma unroll 1
        for (int gg = 0; gg < ng; gg += 2) {
          const bool two = gg + 1 < ng;
          const Ops o = ops_of(rd, gg, 4 * gg + k4 < left);
          const Ops p = ops_of(rd, two ? gg + 1 : gg, two && 4 * (gg + 1) + k4 < left);
          const unsigned m0 = (unsigned)__builtin_amdgcn_readlane(tcur.y, 4 * gg);
          const unsigned m1 = (unsigned)__builtin_amdgcn_readlane(tcur.y, 4 * gg + 4);
          mfmas(o, p, m0 | (two ? m1 : 0u));
        }
      }
      tcur = tnext;
      tnext = tnn;
      // the next issue overwrites this buffer: its reads are done first
      // (LDS operations of a wave complete in order)
      __builtin_amdgcn_wave_barrier();
    };
    for (int b = 0; b < nbatches; b += 2) {
      half(b, rb0, rb1);
      half(b + 1, rb1, rb0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): no load left in flight
  }

  // write-out, one x plane at a time.  Lane (k4, n16) holds, for block
  // (u, p = 2 py + pz), the cells (y, z) = (4 py + k4, 4 pz + r), r = 0..3,
  // of plane n16.
  const bool shared = item.w != 0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int gx = origin[0] + u;
    if (gx >= W.n[0]) break;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int py = p >> 1, pz = p & 1;
      *reinterpret_cast<float4*>(stg + n16 * WG_PS + (4 * py + k4) * 8 + 4 * pz) =
          make_float4(acc[4 * u + p][0], acc[4 * u + p][1], acc[4 * u + p][2], acc[4 * u + p][3]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // The plane's output, 1024 floats, in runs of one (channel, y row): 8
    // cells along z, complex (re, im) interleaved (16 floats) or real (8).
    // Consecutive lanes take consecutive floats of a run, so one store or
    // atomic instruction covers whole 64-byte runs.
    constexpr int RL = CPLX ? 16 : 8;  // floats per run
    const int nz = W.n[2] - origin[2] < 8 ? W.n[2] - origin[2] : 8;
    auto run_out = [&](int run) -> float* {  // output of float 0 of a run, or null
      const int c = run >> 3, y = run & 7;
      const int gy = origin[1] + y;
      if (gy >= W.n[1] || c0 + c >= W.nbatch) return nullptr;
      const int64_t cell0 = ((int64_t)gx * W.n[1] + gy) * W.n[2] + origin[2];
      return out + ((c0 + c) * W.ncells + cell0) * (CPLX ? 2 : 1);
    };
    auto stage_at = [&](int run, int w) -> float {  // float w of a run
      const int c = run >> 3, y = run & 7;
      return CPLX ? stg[(2 * c + (w & 1)) * WG_PS + 8 * y + (w >> 1)]
                  : stg[c * WG_PS + 8 * y + w];
    };
    if (shared) {
#pragma unroll 4
      for (int i = 0; i < 16; ++i) {
        const int f = 64 * i + lane, run = f / RL, w = f % RL;
        const float v = stage_at(run, w);
        float* o = run_out(run);
        if (o && (CPLX ? (w >> 1) : w) < nz) unsafeAtomicAdd(o + w, v);
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int f = 4 * (64 * i + lane), run = f / RL, w = f % RL;
        const float4 v = make_float4(stage_at(run, w), stage_at(run, w + 1), stage_at(run, w + 2),
                                     stage_at(run, w + 3));
        float* o = run_out(run);
        if (!o) continue;
        if (nz == 8) {
          *reinterpret_cast<float4*>(o + w) = v;
        } else {
          const float vv[4] = {v.x, v.y, v.z, v.w};
          for (int e = 0; e < 4; ++e)
            if ((CPLX ? ((w + e) >> 1) : w + e) < nz) o[w + e] = vv[e];
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

}  // namespace

bool grid_wave_ok(const tnl_plan_layout& L) {
  if (L.ndim != 3 || L.grid_accum != TNL_GRID_ACCUM_MFMA || L.off_tpm == 0) return false;
  for (int d = 0; d < 3; ++d) {
    if (L.g_tile[d] != 8 || L.ntaps[d] > 5) return false;
    if (L.grid_size[d] < 16 || L.grid_size[d] > 1024) return false;
  }
  return true;
}

int64_t grid_wave_wrec_bytes(const tnl_plan_layout& L) {
  return L.npts * WG_REC * (int64_t)sizeof(float);
}

int grid_wave_launch(const float* vals, const float* locs, float* out, int64_t nbatch,
                     const InterpGeom& G, const tnl_plan_layout& L, const char* ws,
                     const float* rec, float* wrec, int is_complex, hipStream_t s) {
  // wrec: the scratch's window records, unless the plan holds them
  WaveGeom W;
  for (int d = 0; d < 3; ++d) {
    W.n[d] = (int)G.n[d];
    W.ntiles[d] = L.g_ntiles[d];
  }
  W.ncells = G.ncells;
  W.npts = L.npts;
  W.nbatch = nbatch;
  const int64_t ch = is_complex ? WG_NPL / 2 : WG_NPL;
  W.nch = (int)((nbatch + ch - 1) / ch);
  W.circ = G.pad == TNL_PAD_CIRCULAR;
  if (L.off_wrec && L.wrec_kernel == G.kernel && L.wrec_beta == G.beta) {
    wrec = (float*)(ws + L.off_wrec);  // filled by tnl_plan_windows
  } else {
    const int64_t nb = (L.npts + 255) / 256;
    window_rec_kernel<<<(unsigned)(nb < 16384 ? nb : 16384), 256, 0, s>>>(
        locs, L.npts, G, reinterpret_cast<float4*>(wrec));
    TNL_LAUNCH_CHECK("window_rec");
  }
  const int* nitems = (const int*)(ws + L.off_itemoff) + L.g_ntiles_total;
  const int4* items = (const int4*)(ws + L.off_items);
  const int2* tpm = (const int2*)(ws + L.off_tpm);
  const int64_t span = WG_XCD * WG_RUN;
  const int64_t nwg = (L.item_capacity * W.nch + span - 1) / span * span;
  if (nwg > 0x7fffffff) return fail(TNL_EINVAL, "too many gridding work items");
  if (!rec) return fail(TNL_EINVAL, "wave gridding needs the point-major records");
  if (is_complex)
    grid_wave_kernel<true><<<(unsigned)nwg, 64, 0, s>>>(out, W, tpm, wrec, rec, items, nitems);
  else
    grid_wave_kernel<false><<<(unsigned)nwg, 64, 0, s>>>(out, W, tpm, wrec, rec, items, nitems);
  TNL_LAUNCH_CHECK("grid_wave");
  return TNL_OK;
}

}  // namespace tnl

using namespace tnl;

extern "C" int tnl_plan_windows(tnl_plan_layout* layout, void* workspace, const float* locs,
                                int kernel, float beta, void* stream) {
  if (!layout) return fail(TNL_EINVAL, "null layout");
  tnl_plan_layout& L = *layout;
  if (!L.off_wrec) return TNL_OK;  // no window section in this plan
  if (!workspace) return fail(TNL_EINVAL, "null workspace");
  if (L.npts > 0 && !locs) return fail(TNL_EINVAL, "null locs");
  if (kernel != TNL_KERNEL_KAISER_BESSEL && kernel != TNL_KERNEL_SPLINE)
    return fail(TNL_EINVAL, "Unrecognized kernel");
  // the geometry of the plan's axes (its plans are 3-D)
  InterpGeom G;
  std::memset(&G, 0, sizeof(G));
  G.ndim = L.ndim;
  G.kernel = kernel;
  G.norm = 1;
  G.pad = L.pad_mode;
  G.beta = beta;
  for (int d = 0; d < 3; ++d) {
    G.n[d] = L.grid_size[d];
    G.half[d] = L.width[d] * 0.5f;
    G.inv_half[d] = 1.0f / (L.width[d] * 0.5f);
    G.ntap[d] = L.ntaps[d];
  }
  G.ncells = G.n[0] * G.n[1] * G.n[2];
  L.wrec_kernel = -1;  // invalid while being rewritten
  if (L.npts > 0) {
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = (L.npts + 255) / 256;
    window_rec_kernel<<<(unsigned)(nb < 16384 ? nb : 16384), 256, 0, s>>>(
        locs, L.npts, G, reinterpret_cast<float4*>((char*)workspace + L.off_wrec));
    TNL_LAUNCH_CHECK("plan windows");
  }
  L.wrec_kernel = kernel;
  L.wrec_beta = beta;
  return TNL_OK;
}
