/*
 * tnl.h -- C ABI of the MI355X gridding library (libtnl.so).
 *
 * Plain pointers and sizes only; every device pointer is an HBM address
 * owned by the caller; every call is asynchronous on the HIP stream passed
 * as `stream` (a hipStream_t, NULL = the null stream).  Calls are reentrant:
 * the library keeps no mutable global state besides a thread-local error
 * string.  Every entry point returns TNL_OK (0) or a TNL_E* status; the
 * message for the last failure on the calling thread is tnl_last_error().
 *
 * Complex data is passed as interleaved (re, im) float32, exactly as the
 * reference passes torch.view_as_real(...) to its Triton kernels.
 */
#ifndef TNL_H
#define TNL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define TNL_OK 0
#define TNL_EINVAL 1      /* bad argument (maps to ValueError)             */
#define TNL_ESHAPE 2      /* shape mismatch (maps to RuntimeError)         */
#define TNL_EHIP 3        /* HIP runtime error (launch / memset)           */
#define TNL_EWORKSPACE 4  /* workspace too small                           */

/* ---- enums mirrored from the reference's string arguments ------------ */
#define TNL_KERNEL_KAISER_BESSEL 0 /* kernel="kaiser_bessel"             */
#define TNL_KERNEL_SPLINE 1        /* kernel="spline"                    */
#define TNL_PAD_CIRCULAR 0         /* pad_mode="circular"                */
#define TNL_PAD_ZERO 1             /* pad_mode="zero"                    */
#define TNL_F32 0
#define TNL_F64 1
#define TNL_I32 2  /* integer element types: tnl_index_adjoint only */
#define TNL_I64 3
#define TNL_I16 4
#define TNL_I8 5
#define TNL_U8 6

/* Largest per-dimension tap count ceil(w)+1 the kernels accept.  The
 * reference rejects next_pow2(ceil(w+1)) > 64 (doubled for complex data)
 * in get_block_width (src/torchlinops/functional/_interp/grid.py:145-158). */
#define TNL_MAX_TAPS 64

const char* tnl_last_error(void);
int tnl_abi_version(void);

/* ---- point-binning plan --------------------------------------------------
 * A plan sorts the nonuniform points of one trajectory into spatial tiles
 * so that grid/ungrid read and write HBM tile by tile.  It depends only on
 * (locs, ndim, grid_size, width, pad_mode), never on values, so one plan
 * serves every coil, every CG iteration and both directions.
 * No reference counterpart: the reference's Triton kernels walk points in
 * input order (src/torchlinops/functional/_interp/ungrid.py:154-503).
 */
/* Tuning knobs for the plan (0 = library default).  The autotuner in the
 * Python package searches these; they are copied into the layout, so a
 * plan always runs with the tiling it was built for. */
typedef struct tnl_tuning {
  int32_t u_tile[3];  /* interpolation bin edge per (left-padded) axis      */
  int32_t g_tile[3];  /* gridding tile edge per (left-padded) axis          */
  int32_t max_sub;    /* points per interpolation subproblem                */
  int32_t g_cap;      /* touches per gridding work item                     */
  int32_t lds_budget; /* LDS bytes per workgroup for channel chunking       */
  int32_t grid_accum; /* TNL_GRID_ACCUM_*: gridding accumulator           */
} tnl_tuning;

/* Gridding accumulator.  AUTO takes the matrix-core (MFMA) accumulator for
 * 3-D plans with 8x8x8 and 2-D plans with 16x32 gridding tiles (the
 * defaults for grids above 16 / 32 x 64) and the LDS accumulator
 * otherwise; MFMA on a plan that does not qualify is an error. */
#define TNL_GRID_ACCUM_AUTO 0
#define TNL_GRID_ACCUM_LDS 1
#define TNL_GRID_ACCUM_MFMA 2

typedef struct tnl_plan_layout {
  int32_t ndim;
  int32_t ntaps[3];      /* ceil(w_d)+1 per dim (dims left-padded to 3)      */
  /* interpolation bins: each point goes to the bin holding its first tap   */
  int32_t u_tile[3];     /* bin edge per dim (cells)                         */
  int32_t u_ntiles[3];
  int64_t u_nbins;       /* prod(u_ntiles); bin u_nbins holds dead points    */
  /* gridding tiles: each tile lists every point whose taps touch it      */
  int32_t g_tile[3];
  int32_t g_ntiles[3];
  int64_t g_ntiles_total;
  int64_t grid_size[3];  /* left-padded                                       */
  float width[3];        /* left-padded                                       */
  int32_t pad_mode;
  int64_t npts;
  int64_t max_sub;       /* points per interpolation subproblem              */
  int64_t sub_capacity;  /* upper bound on interpolation subproblems         */
  int64_t g_cap;         /* touches per gridding work item                   */
  int64_t touch_capacity;/* upper bound on (tile, point) touches             */
  int64_t item_capacity; /* upper bound on gridding work items               */
  int64_t lds_budget;    /* LDS bytes per workgroup                          */
  int64_t grid_accum;    /* TNL_GRID_ACCUM_* (resolved: never AUTO)          */
  /* byte offsets inside the workspace                                      */
  int64_t off_counts, off_starts, off_cursor, off_perm, off_slocs;
  int64_t off_suboff, off_sub;
  int64_t off_gcounts, off_gstarts, off_gcursor, off_tlist;
  int64_t off_itemoff, off_items;
  int64_t off_items_tmp; /* staging for the item ordering during the build
                          * (3-D plans list their work items largest first)  */
  int64_t off_tpm;       /* 3-D matrix-core plans (8^3 tiles, <= 5 taps, every
                          * axis 16..1024 cells): per touch (item order) the
                          * caller index of the point and the union of the
                          * tile's MFMA blocks that its group of 4 touches
                          * (counted from the item's first) covers, as int32
                          * pairs; 0 when absent                            */
  int64_t off_wrec;      /* plans with touch pairs: one 64-byte window record
                          * per point (tap weights per axis, first taps),
                          * filled by tnl_plan_windows; 0 when absent        */
  int32_t wrec_kernel;   /* kernel and beta the window records were last    */
  float wrec_beta;       /* filled for; wrec_kernel = -1 until then          */
  int64_t bytes;        /* total workspace bytes                             */
} tnl_plan_layout;

/* Fill *out for a trajectory of npts points; tuning may be NULL. */
int tnl_plan_layout_query(int64_t npts, int ndim, const int64_t* grid_size,
                          const float* width, int pad_mode,
                          const tnl_tuning* tuning, tnl_plan_layout* out);

/* Build the plan described by *layout into a caller-owned device workspace
 * of at least layout->bytes. */
int tnl_plan_build(const float* locs, const tnl_plan_layout* layout,
                   void* workspace, int64_t workspace_bytes, void* stream);

/* Fill the window records of a plan with touch pairs (layout->off_wrec !=
 * 0) for one kernel and beta, from the same locs the plan was built from,
 * and record them in layout->wrec_kernel / wrec_beta.  Gridding calls with
 * that plan, kernel and beta then read the records from the plan instead
 * of evaluating them into their scratch on every call.  A plan without the
 * section is left as it is.  No counterpart in the reference (its kernels
 * evaluate the window per tap, src/torchlinops/functional/_interp/
 * grid.py:161-513). */
int tnl_plan_windows(tnl_plan_layout* layout, void* workspace, const float* locs,
                     int kernel, float beta, void* stream);

/* ---- interpolation (ungrid) -------------------------------------------
 * out[b, p] = sum_{g in taps(p)} w(t_p - g) * vals[b, g]
 * Replaces UNGRID[ndim][grid](vals, locs, output, nbatch, npts, kernel,
 * norm, pad_mode, is_complex, *grid_size, *width, *BLOCK_WIDTH, beta)
 * at src/torchlinops/functional/_interp/ungrid.py:103-117.
 * vals: [nbatch, *grid_size] (x2 floats if complex); locs: [npts, ndim];
 * out: [nbatch, npts] (x2 if complex), fully written by the call.
 * layout/plan: both NULL, or the layout and workspace of a plan built by
 * tnl_plan_build for the same locs, grid_size, width and pad_mode.
 */
int tnl_ungrid(const float* vals, const float* locs, float* out,
               int64_t nbatch, int64_t npts, int ndim,
               const int64_t* grid_size, const float* width, int kernel,
               int norm, int pad_mode, int is_complex, float beta,
               const tnl_plan_layout* layout, const void* plan, void* stream);

/* Scratch bytes that tnl_ungrid_ws uses: npts x nbatch channel values
 * (x2 floats if complex). */
int tnl_ungrid_scratch_bytes(int64_t npts, int64_t nbatch, int is_complex, int64_t* bytes);

/* tnl_ungrid with a caller-owned, 16-byte aligned device scratch of at
 * least tnl_ungrid_scratch_bytes bytes (contents on entry ignored and
 * overwritten).  With a plan, the interpolation writes each point's values
 * of all channels side by side into the scratch (whole 64-byte lines for
 * 8 complex channels) and one tiled transpose writes them to `out` in
 * its [nbatch, npts] layout.  For trajectories listed in scattered order
 * (neighbouring points in different bins) this replaces the channel-major
 * stores, which leave each 64-byte line of `out` partly written by many
 * workgroups; for trajectories in acquisition order (spokes, spirals) the
 * plain call is faster.  Without a plan the scratch is unused.  No
 * counterpart in the reference (it stands beside ungrid.py:103-117).  Same
 * results as tnl_ungrid, bit for bit. */
int tnl_ungrid_ws(const float* vals, const float* locs, float* out, int64_t nbatch,
                  int64_t npts, int ndim, const int64_t* grid_size, const float* width,
                  int kernel, int norm, int pad_mode, int is_complex, float beta,
                  const tnl_plan_layout* layout, const void* plan, void* scratch,
                  int64_t scratch_bytes, void* stream);

/* ---- gridding (interpolate adjoint) ------------------------------------
 * out[b, g] = sum_{p : g in taps(p)} w(t_p - g) * vals[b, p]
 * Replaces GRID[ndim][grid](...) at
 * src/torchlinops/functional/_interp/grid.py:107-122.  The reference's
 * caller zero-fills `out` first (grid.py:96-101); here the call writes
 * every element of out itself, so the caller may pass uninitialised memory.
 */
int tnl_grid(const float* vals, const float* locs, float* out, int64_t nbatch,
             int64_t npts, int ndim, const int64_t* grid_size,
             const float* width, int kernel, int norm, int pad_mode,
             int is_complex, float beta, const tnl_plan_layout* layout,
             const void* plan, void* stream);

/* ---- float64 / complex128 interpolation and gridding --------------------
 * tnl_ungrid / tnl_grid for double data (vals, locs and out as doubles,
 * complex interleaved), computed in double precision with the windows
 * evaluated in double; no plan (one thread per point; gridding adds with
 * double atomics into out, which the call zero-fills first).  Stand beside
 * the same dispatch points (ungrid.py:86-119, grid.py:91-124): the
 * reference's torch fallback computes in the input's dtype
 * (grid.py:571-639, ungrid.py:624-684). */
int tnl_ungrid_f64(const double* vals, const double* locs, double* out, int64_t nbatch,
                   int64_t npts, int ndim, const int64_t* grid_size, const float* width,
                   int kernel, int norm, int pad_mode, int is_complex, double beta,
                   void* stream);
int tnl_grid_f64(const double* vals, const double* locs, double* out, int64_t nbatch,
                 int64_t npts, int ndim, const int64_t* grid_size, const float* width,
                 int kernel, int norm, int pad_mode, int is_complex, double beta,
                 void* stream);

/* Scratch bytes that tnl_grid_ws uses with this plan for nbatch channels:
 * the channel values re-packed point-major, one 64-byte record (up to 8
 * complex / 16 real channel values) per point and channel chunk, for the
 * matrix-core gridding kernels (plans whose grid_accum is
 * TNL_GRID_ACCUM_MFMA), plus, for plans with touch pairs (off_tpm != 0),
 * one 64-byte window record per point (its tap weights per axis and first
 * taps, evaluated once per call instead of once per touched tile); 0 for
 * other plans (tnl_grid_ws then equals tnl_grid). */
int tnl_grid_scratch_bytes(const tnl_plan_layout* layout, int64_t nbatch,
                           int is_complex, int64_t* bytes);

/* tnl_grid with a caller-owned device scratch buffer of at least
 * tnl_grid_scratch_bytes bytes (its contents on entry are ignored and
 * overwritten).  With the records each touch of a gridding tile reads one
 * contiguous record instead of its channel values scattered over the
 * coil-major input.  No counterpart in the reference: its Triton kernel
 * reads the channel values of each point where they lie
 * (src/torchlinops/functional/_interp/grid.py:161-513).  Same results as
 * tnl_grid, bit for bit, except for 3-D plans with touch pairs (off_tpm),
 * where one wave accumulates a whole tile in its own order: equal to
 * float32 rounding. */
int tnl_grid_ws(const float* vals, const float* locs, float* out, int64_t nbatch,
                int64_t npts, int ndim, const int64_t* grid_size,
                const float* width, int kernel, int norm, int pad_mode,
                int is_complex, float beta, const tnl_plan_layout* layout,
                const void* plan, void* scratch, int64_t scratch_bytes,
                void* stream);

/* Largest scratch (bytes) tnl_grid_static can use: 64 MiB, i.e. 1 048 576
 * 64-byte records (points x channel chunks of 8 complex / 16 real, plus one
 * window record per point for plans with touch pairs). */
#define TNL_GRID_STATIC_SCRATCH_BYTES (64ll << 20)

/* tnl_grid_ws without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): the point records live in a file-scope
 * device array of TNL_GRID_STATIC_SCRATCH_BYTES bytes that each GPU holds
 * from the library's load to the end of the process.  No counterpart in the
 * reference.  Plans needing more scratch than that (tnl_grid_scratch_bytes)
 * are rejected with TNL_EINVAL before any device work.  The call allocates
 * nothing; every call on one GPU shares that GPU's array (its contents are
 * rewritten by each call, nothing needs to start at zero), so calls that
 * overlap on one GPU (two streams, two host threads) race and must be
 * serialised by the caller.  Same arguments as tnl_grid and the same
 * results as tnl_grid_ws. */
int tnl_grid_static(const float* vals, const float* locs, float* out, int64_t nbatch,
                    int64_t npts, int ndim, const int64_t* grid_size,
                    const float* width, int kernel, int norm, int pad_mode,
                    int is_complex, float beta, const tnl_plan_layout* layout,
                    const void* plan, void* stream);

/* ---- many problems in one persistent launch ----------------------------
 * A batch is a set of independent interpolation problems of one ndim,
 * kernel, norm, pad_mode and dtype, each with its own grid size, width,
 * beta, channel count, trajectory and plan (tnl_plan_build).  Grid-side
 * data of all problems live
 * in one buffer, point-side data in another, each problem at its own
 * element offset:
 *   grid buffer   problem i: [nbatch_i, *grid_size_i] at grid_offset_i
 *   point buffer  problem i: [nbatch_i, npts_i]       at pts_offset_i
 * tnl_batch_build copies a problem table into a device workspace once;
 * tnl_ungrid_batched / tnl_grid_batched then run every problem's work in
 * one persistent launch each (no host work, capturable in a graph).
 * Results equal tnl_ungrid / tnl_grid called per problem with its plan.
 * No counterpart in the reference, which launches per problem
 * (src/torchlinops/functional/_interp/ungrid.py:86-119, grid.py:91-124).
 */
typedef struct tnl_batch_problem {
  int64_t nbatch;         /* channels (coils) */
  int64_t npts;
  int64_t grid_size[3];   /* first ndim entries */
  float width[3];         /* first ndim entries */
  float beta;             /* Kaiser-Bessel beta */
  int64_t grid_offset;    /* element offset in the grid buffer */
  int64_t pts_offset;     /* element offset in the point buffer */
  const tnl_plan_layout* layout;  /* this problem's plan */
  const void* plan;
} tnl_batch_problem;

typedef struct tnl_batch_layout {
  int32_t nprob, ndim, kernel, norm, pad_mode, is_complex;
  int32_t nt;             /* tap capacity per axis of the launches */
  int32_t ungrid_cc;      /* channels per ungrid workgroup */
  int32_t grid_cc;        /* channels per gridding accumulator */
  int32_t ungrid_chunks;  /* channel chunks an ungrid workgroup walks:
                             max over problems of ceil(nbatch / ungrid_cc) */
  int32_t grid_mfma;      /* 1: every plan's grid_accum is TNL_GRID_ACCUM_MFMA,
                             so the batch grids on the matrix cores */
  int32_t reserved;
  int64_t total_sub;      /* sum of the plans' subproblem capacities */
  int64_t total_items;    /* sum of the plans' gridding work-item capacities */
  int64_t ungrid_lds;     /* LDS bytes per ungrid workgroup */
  int64_t grid_lds;       /* LDS bytes per gridding workgroup */
  int64_t bytes;          /* device workspace for the problem table */
} tnl_batch_layout;

int tnl_batch_layout_query(const tnl_batch_problem* probs, int nprob, int ndim,
                           int kernel, int norm, int pad_mode, int is_complex,
                           tnl_batch_layout* layout);
/* Blocking (synchronises `stream`): call once, outside graph capture. */
int tnl_batch_build(const tnl_batch_problem* probs,
                    const tnl_batch_layout* layout, void* ws, void* stream);
int tnl_ungrid_batched(const float* grid_vals, float* pts_out,
                       const tnl_batch_layout* layout, const void* ws,
                       void* stream);
int tnl_grid_batched(const float* pts_vals, float* grid_out,
                     const tnl_batch_layout* layout, const void* ws,
                     void* stream);

/* ---- segmented solver kernels -------------------------------------------
 * For conjugate gradients over many independent problems held in one flat
 * float buffer (complex as interleaved floats).  Segment s is the float
 * range of problem s; seg_len[s] is its length in floats.  Dot products
 * are Re(conj(a) . b) per segment, accumulated in float64 in a fixed order
 * (deterministic, no atomics, no memsets: capturable in a graph).  The CG
 * step follows src/torchlinops/alg/pcg.py:67-87 per segment:
 *   tnl_seg_cg_update:    alpha = rs / pAp;  x += alpha p;  r -= alpha Ap;
 *                         rs_new = |r|^2
 *   tnl_seg_cg_direction: beta = rs_new / rs; p = r + beta p; rs = rs_new
 * (alpha, beta are 0 where their denominator is 0).  No counterpart in the
 * reference, whose CG (pcg.py:13-88) runs one problem with host scalars.
 */
typedef struct tnl_seg_layout {
  int32_t nseg;
  int32_t reserved;
  int64_t numel;         /* floats in the buffer */
  int64_t block_floats;  /* floats per reduction block */
  int64_t nblocks;
  int64_t off_blocks, off_blkseg, off_segblk, off_partial;
  int64_t bytes;         /* device workspace */
} tnl_seg_layout;

int tnl_seg_layout_query(const int64_t* seg_len, int nseg, int64_t block_floats,
                         tnl_seg_layout* layout);
/* Blocking (synchronises `stream`): call once, outside graph capture. */
int tnl_seg_build(const int64_t* seg_len, const tnl_seg_layout* layout, void* ws,
                  void* stream);
int tnl_seg_dot(const float* a, const float* b, double* out,
                const tnl_seg_layout* layout, void* ws, void* stream);
int tnl_seg_cg_update(float* x, float* r, const float* p, const float* Ap,
                      const double* rs, const double* pAp, double* rs_new,
                      const tnl_seg_layout* layout, void* ws, void* stream);
int tnl_seg_cg_direction(float* p, const float* r, double* rs,
                         const double* rs_new, const tnl_seg_layout* layout,
                         void* ws, void* stream);

/* ---- 3x3x3 convolutions of the learned preconditioner -------------------
 * Channel-last bf16 activations [nbatch, z, y, x, c_p] (c_p a multiple of
 * 8, padding channels zero); weights bf16 [16][Kp], row = output channel,
 * column tap * cin_p + ci, tap = (dz+1)*9 + (dy+1)*3 + (dx+1), Kp =
 * 32 * ceil(27 cin_p / 32), rows >= cout and columns >= 27 cin_p zero.
 * "Same" zero padding; fp32 accumulation on the matrix cores (BF16 MFMA).
 *   tnl_conv3d_fwd: out = [relu](conv(in, W) + bias), written as bf16
 *     and/or fp32 [nbatch, z, y, x, cout_p] (either pointer may be null;
 *     channels >= cout are not written).  cin_p in {8, 16}, cout <= 16.
 *     The data gradient is this call with flipped, transposed weights.
 *   tnl_conv3d_wgrad: dw[co][ci][tap] (fp32 [cout][cin_p][27]) = sum over
 *     voxels of dout[co](p) * in[ci](p + tap); ws holds per-tile partials
 *     (tnl_conv3d_wgrad_workspace bytes), summed in a fixed order.
 * No counterpart in the reference (the learned preconditioner is an
 * extension beside its CG, src/torchlinops/alg/pcg.py).
 */
int tnl_conv3d_fwd(const void* in, const void* w, const float* bias,
                   void* out_bf16, float* out_f32, int64_t nbatch, int z, int y,
                   int x, int cin_p, int cout, int cout_p, int relu, void* stream);
int64_t tnl_conv3d_wgrad_workspace(int64_t nbatch, int z, int y, int x);
int tnl_conv3d_wgrad(const void* in, const void* dout, float* dw, float* ws,
                     int64_t ws_bytes, int64_t nbatch, int z, int y, int x,
                     int cin_p, int cout, int cout_p, void* stream);

/* ---- trajectory gradient of interpolate / interpolate_adjoint ----------
 * grad_locs[p, d] = sum_b sum_{g in taps(p)}
 *                     Re(conj(pt_vals[b, p]) * grid_vals[b, g]) * dW(g, t_p)/dt_{p,d}
 * with W the window weight of tnl_ungrid / tnl_grid and its derivative
 * taken analytically (Kaiser-Bessel: -beta^2 u I1(z)/z with the
 * Abramowitz & Stegun 9.8.3/9.8.4 approximations; spline: -1 on the
 * support).  For ungrid (y = A(t) x) pass grid_vals = x, pt_vals = dL/dy;
 * for grid (X = A(t)^H v) pass grid_vals = dL/dX, pt_vals = v.  locs and
 * grad_locs are float32 [npts, ndim]; grad_locs is fully written.
 * No counterpart in the reference: its autograd pair returns None for the
 * locs gradient (src/torchlinops/functional/_interp/interp.py:46-70,
 * 157-181).  Kernel widths up to 15.
 */
int tnl_interp_grad_locs(const float* grid_vals, const float* pt_vals,
                         const float* locs, float* grad_locs, int64_t nbatch,
                         int64_t npts, int ndim, const int64_t* grid_size,
                         const float* width, int kernel, int norm,
                         int pad_mode, int is_complex, float beta,
                         void* stream);

/* ---- array_to_blocks (unfold) -------------------------------------------
 * y[n, b..., j...] = x[n, b*s + j]  (pure copy, bit-exact)
 * Replaces UNFOLD[ndim][grid](x, y, nbatch, *nblocks, *block_size,
 * *im_size, *stride, *BLOCK_SIZE) at
 * src/torchlinops/functional/_unfold/unfold.py:117-126.  Sizes are those
 * after the reference's complex interleave (last dim, block and stride x2,
 * unfold.py:581-587).  elem_bytes in {1, 2, 4, 8}; y is fully written.
 */
int tnl_unfold(const void* x, void* y, int elem_bytes, int64_t nbatch,
               int ndim, const int64_t* nblocks, const int64_t* block_size,
               const int64_t* im_size, const int64_t* stride, void* stream);

/* ---- blocks_to_array (fold) ---------------------------------------------
 * y[n, i] = sum over blocks covering i, in lexicographic block order, of
 * x[n, b, i - b*s]; uncovered cells are 0.  Bit-exact against the
 * reference's sequential `out[slc] += x[blk]` (fold.py:489-517).
 * Replaces FOLD[ndim][grid](x, y, nbatch, *nblocks, *block_size, *im_size,
 * *stride) at src/torchlinops/functional/_unfold/fold.py:105-113.
 * dtype: TNL_F32 or TNL_F64; y is fully written.
 */
int tnl_fold(const void* x, void* y, int dtype, int64_t nbatch, int ndim,
             const int64_t* nblocks, const int64_t* block_size,
             const int64_t* im_size, const int64_t* stride, void* stream);

/* ---- index (integer sampling gather) ------------------------------------
 * out[n, p] = x[n, idx[p]] over a flattened grid x [nbatch, ngrid] and
 * raveled cell indices idx [npts] (int64, in [0, ngrid)); pure copy,
 * bit-exact.  Replaces torch.index_select(x_flat, -1, idx_flat) in
 * multi_index (src/torchlinops/functional/_index/index.py:52-72), which
 * index (index.py:17-49) and Sampling.fn (src/torchlinops/linops/
 * sampling.py:96-98) call.  elem_bytes in {1, 2, 4, 8, 16} (complex
 * values move as one element).  An index outside [0, ngrid) reads as
 * zero; callers check ranges before the call.  out is fully written.
 */
int tnl_index(const void* x, const int64_t* idx, void* out, int elem_bytes,
              int64_t nbatch, int64_t ngrid, int64_t npts, void* stream);

/* ---- index_adjoint (integer sampling scatter-add) -----------------------
 * out[n, c] = sum over points p with idx[p] == c of x[n, p]; cells no
 * point hits are 0.  Replaces y.index_add_(-1, idx_flat, x_flat) in
 * multi_grid (src/torchlinops/functional/_index/index.py:140-180), which
 * index_adjoint (index.py:75-99) and Sampling.adj_fn (sampling.py:100-101)
 * call.  The points are handed over grouped by cell: order [npts] lists
 * the point numbers sorted by cell (stable, so ascending within a cell)
 * and starts [ngrid + 1] is the offset of each cell's run in order.
 * A cell with at most heavy_min points (every cell when nheavy == 0) is
 * summed in ascending p from zero: bit-exact against the CPU's sequential
 * `y[idx[p]] += x[p]`.  Cells with more points must all be listed in
 * heavy [nheavy]; each is summed by one workgroup in a fixed order
 * (deterministic, within rounding of the sequential sum).  dtype TNL_F32
 * or TNL_F64, is_complex for interleaved (re, im) pairs; or an integer type
 * (TNL_I8, TNL_U8, TNL_I16, TNL_I32, TNL_I64; real only), summed exactly and
 * wrapping like the reference's index_add_ of that type.  out is fully
 * written (no zeroing needed, no atomics).
 */
int tnl_index_adjoint(const void* x, const int64_t* order, const int64_t* starts,
                      const int64_t* heavy, int64_t nheavy, int64_t heavy_min,
                      void* out, int dtype, int is_complex, int64_t nbatch,
                      int64_t ngrid, int64_t npts, void* stream);

/* ---- Toeplitz normal: centred pad + ifftshift, fftshift + centred crop ---
 * Image index i_d in [0, N_d) has centred coordinate c_d = i_d - N_d/2.
 * tnl_pad_wrapped:  y[n, q] = x[n, i] where q_d = c_d mod P_d on every
 *                   axis, 0 at every other cell of y [nbatch, *pad_size];
 *                   equals ifftshift(Pad(x)) of the reference's centred pad
 *                   (src/torchlinops/linops/pad_last.py:139-199) and FFT's
 *                   ifftshift (src/torchlinops/utils/_fft.py:7-31).
 * tnl_crop_wrapped: x[n, i] = y[n, (c mod P)]; equals Crop(fftshift(y)).
 * Together they replace the four centring rolls and the pad/crop copies
 * of the Toeplitz normal Crop . F^H . (K *) . F . Pad built at
 * src/torchlinops/linops/nufft.py:226-239, the kernel K then being taken
 * in natural FFT order.  weight: NULL (pure copies, bit-exact;
 * elem_bytes in {4, 8, 16}) or float32 [*im_size] multiplying every copied
 * value (float32 / complex64 data, elem_bytes 4 or 8): the NUFFT's
 * Apodize and 1/scale (nufft.py:166-184) folded into the same pass, so
 * that NUFFT = interp(fftn(pad_wrapped(x, apod/scale))) at locations
 * relabelled by -pad_size/2 and NUFFT^H its mirror.
 * 1 <= im_size <= pad_size; outputs fully written.
 */
int tnl_pad_wrapped(const void* x, void* y, int elem_bytes, int64_t nbatch,
                    int ndim, const int64_t* im_size, const int64_t* pad_size,
                    const float* weight, void* stream);
int tnl_crop_wrapped(const void* y, void* x, int elem_bytes, int64_t nbatch,
                     int ndim, const int64_t* im_size, const int64_t* pad_size,
                     const float* weight, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TNL_H */
