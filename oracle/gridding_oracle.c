/*
 * gridding_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's gridding path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * Nothing in the product links or calls this file.
 *
 * Follows (paths relative to the reference repository):
 *   kaiser_bessel_torch      src/torchlinops/functional/_interp/kernels.py:28-89
 *   spline_torch             kernels.py:92-94
 *   tap set / weights        kernels.py:108-147 (weights_torch) with the
 *                            Triton keep-test |t - g| <= w/2 of
 *                            kernels.py:264-279, 299-302, 349-354
 *   ungrid_torch             src/torchlinops/functional/_interp/ungrid.py:624-684
 *   grid_torch               src/torchlinops/functional/_interp/grid.py:571-639
 *   _unfold_torch            src/torchlinops/functional/_unfold/unfold.py:531-561
 *   _fold_torch              src/torchlinops/functional/_unfold/fold.py:489-517
 *   orc_grad_locs_f64        no reference counterpart (the reference returns
 *                            no locs gradient, interp.py:46-70, 157-181): the
 *                            analytic d/dt of the weights above, pinned by
 *                            scipy's I1 and by float64 finite differences
 *
 * Order of operations follows the reference: taps enumerated as
 * meshgrid(arange(w+1), indexing="ij") (x slowest), per-axis window values
 * multiplied x, y, z in turn (.prod(dim=-1)), gridding accumulated
 * sequentially batch -> point -> tap (index_put_(accumulate=True) on CPU),
 * fold accumulated block by block in lexicographic order.  The float32
 * functions keep every intermediate in float32 (build with
 * -ffp-contract=off so no FMA contraction changes rounding); the float64
 * functions are the same algorithm in double precision.
 *
 * Known deviation (documented, deliberate): the reference torch fallback
 * masks taps with |t - g| / (w/2) <= w/2 (kernels.py:126,136), which for
 * w < 2 is stricter than the Triton / sigpy rule |t - g| <= w/2 that the
 * GPU path follows; this file follows the GPU rule.  Norm 2 on 2-D/3-D is
 * an error on the reference torch path and yields NaN in the reference
 * Triton path for Kaiser-Bessel taps outside the unit ball; here such taps
 * get weight 0.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define KB 0
#define SPLINE 1
#define CIRCULAR 0
#define ZERO 1

/* ---------------------------------------------------------------- KB */
/* Same evaluation order as kaiser_bessel_torch (kernels.py:50-89). */
float orc_kb_f32(float x, float beta) {
  float z = beta * sqrtf(1.0f - x * x);
  int small = z < 3.75f;
  float t = z / 3.75f;
  float t2 = t * t, t4 = t2 * t2, t6 = t4 * t2, t8 = t4 * t4, t10 = t6 * t4,
        t12 = t6 * t6;
  float t3 = t2 * t, t5 = t3 * t2, t7 = t5 * t2;
  if (small) {
    return 1.0f + 3.5156229f * t2 + 3.0899424f * t4 + 1.2067492f * t6 +
           0.2659732f * t8 + 0.0360768f * t10 + 0.0045813f * t12;
  }
  return expf(z) / sqrtf(z) *
         (0.39894228f + 0.01328592f / t + 0.00225319f / t2 -
          0.00157565f / t3 + 0.00916281f / t4 - 0.02057706f / t5 +
          0.02635537f / t6 - 0.01647633f / t7 + 0.00392377f / t8);
}

double orc_kb_f64(double x, double beta) {
  double z = beta * sqrt(1.0 - x * x);
  double t = z / 3.75;
  if (z < 3.75) {
    double t2 = t * t;
    return 1.0 + 3.5156229 * t2 + 3.0899424 * pow(t2, 2) +
           1.2067492 * pow(t2, 3) + 0.2659732 * pow(t2, 4) +
           0.0360768 * pow(t2, 5) + 0.0045813 * pow(t2, 6);
  }
  double r = 1.0 / t;
  return exp(z) / sqrt(z) *
         (0.39894228 + 0.01328592 * r + 0.00225319 * pow(r, 2) -
          0.00157565 * pow(r, 3) + 0.00916281 * pow(r, 4) -
          0.02057706 * pow(r, 5) + 0.02635537 * pow(r, 6) -
          0.01647633 * pow(r, 7) + 0.00392377 * pow(r, 8));
}

void orc_kb_vec_f32(const float* x, float* out, int64_t n, float beta) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_kb_f32(x[i], beta);
}
void orc_kb_vec_f64(const double* x, double* out, int64_t n, double beta) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_kb_f64(x[i], beta);
}

/* d/du I0(beta sqrt(1 - u^2)) = -beta^2 u I1(z)/z, z = beta sqrt(1 - u^2),
 * with I1(z)/z from Abramowitz & Stegun 9.8.3 (z < 3.75) and 9.8.4. */
double orc_kb_deriv_f64(double u, double beta) {
  double z = beta * sqrt(1.0 - u * u);
  double t = z / 3.75, g;
  if (z < 3.75) {
    double t2 = t * t;
    g = 0.5 + 0.87890594 * t2 + 0.51498869 * pow(t2, 2) + 0.15084934 * pow(t2, 3) +
        0.02658733 * pow(t2, 4) + 0.00301532 * pow(t2, 5) + 0.00032411 * pow(t2, 6);
  } else {
    double r = 1.0 / t;
    g = exp(z) / (z * sqrt(z)) *
        (0.39894228 - 0.03988024 * r - 0.00362018 * pow(r, 2) +
         0.00163801 * pow(r, 3) - 0.01031555 * pow(r, 4) +
         0.02282967 * pow(r, 5) - 0.02895312 * pow(r, 6) +
         0.01787654 * pow(r, 7) - 0.00420059 * pow(r, 8));
  }
  return -beta * beta * u * g;
}
void orc_kb_deriv_vec_f64(const double* u, double* out, int64_t n, double beta) {
  for (int64_t i = 0; i < n; ++i) out[i] = orc_kb_deriv_f64(u[i], beta);
}

static float win_f32(int kernel, float d, float beta) {
  if (kernel == KB) return orc_kb_f32(d, beta);
  float a = 1.0f - fabsf(d);
  return a > 0.0f ? a : 0.0f;
}
static double win_f64(int kernel, double d, double beta) {
  if (kernel == KB) return orc_kb_f64(d, beta);
  double a = 1.0 - fabs(d);
  return a > 0.0 ? a : 0.0;
}

static int64_t mod_pos(int64_t g, int64_t n) {
  int64_t r = g % n;
  return r < 0 ? r + n : r;
}

/* ------------------------------------------------------- tap weights
 * For point p: ntap_d = ceil(w_d) + 1 taps per axis, lower_d =
 * ceil(t_d - w_d/2), lexicographic over (k_0, ..., k_{D-1}).  For each tap,
 * writes the flat cell index (or -1 when the tap is dropped) and weight. */
typedef struct {
  int ndim;
  int64_t n[3];
  int ntap[3];
  int64_t ntaps_total;
  int kernel, norm, pad;
} geom_t;

static void make_geom(geom_t* g, int ndim, const int64_t* grid,
                      const float* width, int kernel, int norm, int pad) {
  g->ndim = ndim;
  g->ntaps_total = 1;
  for (int d = 0; d < ndim; ++d) {
    g->n[d] = grid[d];
    g->ntap[d] = (int)ceil((double)width[d]) + 1;
    g->ntaps_total *= g->ntap[d];
  }
  g->kernel = kernel;
  g->norm = norm;
  g->pad = pad;
}

static void taps_f32(const geom_t* G, const float* t, const float* width,
                     float beta, int64_t* cell, float* w) {
  int64_t lower[3];
  float r[3];
  for (int d = 0; d < G->ndim; ++d) {
    r[d] = width[d] / 2.0f;
    lower[d] = (int64_t)ceilf(t[d] - r[d]);
  }
  for (int64_t j = 0; j < G->ntaps_total; ++j) {
    int64_t rem = j, k[3];
    for (int d = G->ndim - 1; d >= 0; --d) {
      k[d] = rem % G->ntap[d];
      rem /= G->ntap[d];
    }
    int keep = 1;
    float wt = 1.0f, s2 = 0.0f;
    int64_t c = 0;
    for (int d = 0; d < G->ndim; ++d) {
      int64_t g = lower[d] + k[d];
      float diff = t[d] - (float)g;
      float dn = diff / r[d];
      if (!(fabsf(diff) <= r[d])) keep = 0;
      if (G->norm == 1)
        wt = wt * win_f32(G->kernel, dn, beta);
      else
        s2 = s2 + dn * dn;
      if (G->pad == ZERO && (g < 0 || g >= G->n[d])) keep = 0;
      c = c * G->n[d] + mod_pos(g, G->n[d]);
    }
    if (G->norm == 2) {
      float rr = sqrtf(s2);
      wt = rr <= 1.0f ? win_f32(G->kernel, rr, beta) : 0.0f;
    }
    if (!isfinite(wt)) wt = 0.0f;
    cell[j] = keep ? c : -1;
    w[j] = keep ? wt : 0.0f;
  }
}

static void taps_f64(const geom_t* G, const double* t, const float* width,
                     double beta, int64_t* cell, double* w) {
  int64_t lower[3];
  double r[3];
  for (int d = 0; d < G->ndim; ++d) {
    r[d] = (double)width[d] / 2.0;
    lower[d] = (int64_t)ceil(t[d] - r[d]);
  }
  for (int64_t j = 0; j < G->ntaps_total; ++j) {
    int64_t rem = j, k[3];
    for (int d = G->ndim - 1; d >= 0; --d) {
      k[d] = rem % G->ntap[d];
      rem /= G->ntap[d];
    }
    int keep = 1;
    double wt = 1.0, s2 = 0.0;
    int64_t c = 0;
    for (int d = 0; d < G->ndim; ++d) {
      int64_t g = lower[d] + k[d];
      double diff = t[d] - (double)g;
      double dn = diff / r[d];
      if (!(fabs(diff) <= r[d])) keep = 0;
      if (G->norm == 1)
        wt = wt * (fabs(dn) <= 1.0 ? win_f64(G->kernel, dn, beta) : 0.0);
      else
        s2 = s2 + dn * dn;
      if (G->pad == ZERO && (g < 0 || g >= G->n[d])) keep = 0;
      c = c * G->n[d] + mod_pos(g, G->n[d]);
    }
    if (G->norm == 2) {
      double rr = sqrt(s2);
      wt = rr <= 1.0 ? win_f64(G->kernel, rr, beta) : 0.0;
    }
    if (!isfinite(wt)) wt = 0.0;
    cell[j] = keep ? c : -1;
    w[j] = keep ? wt : 0.0;
  }
}

static int64_t ncells_of(const geom_t* G) {
  int64_t n = 1;
  for (int d = 0; d < G->ndim; ++d) n *= G->n[d];
  return n;
}

/* ------------------------------------------------------------- ungrid
 * out[b, p] = sum_taps w * vals[b, cell], taps summed in lexicographic
 * order.  comp = 2 for interleaved complex.  Points are independent, so
 * the point loop may run on several threads (nthreads) without changing
 * any result. */
int orc_ungrid_f32(const float* vals, const float* locs, float* out,
                   int64_t nbatch, int64_t npts, int ndim,
                   const int64_t* grid, const float* width, int kernel,
                   int norm, int pad, int is_complex, float beta,
                   int nthreads) {
  geom_t G;
  make_geom(&G, ndim, grid, width, kernel, norm, pad);
  const int comp = is_complex ? 2 : 1;
  const int64_t nc = ncells_of(&G);
  (void)nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * G.ntaps_total);
    float* w = (float*)malloc(sizeof(float) * G.ntaps_total);
#pragma omp for schedule(static)
    for (int64_t p = 0; p < npts; ++p) {
      taps_f32(&G, locs + p * ndim, width, beta, cell, w);
      for (int64_t b = 0; b < nbatch; ++b) {
        for (int c = 0; c < comp; ++c) {
          float acc = 0.0f;
          for (int64_t j = 0; j < G.ntaps_total; ++j) {
            if (cell[j] < 0) continue;
            acc = acc + w[j] * vals[(b * nc + cell[j]) * comp + c];
          }
          out[(b * npts + p) * comp + c] = acc;
        }
      }
    }
    free(cell);
    free(w);
  }
  return 0;
}

int orc_ungrid_f64(const double* vals, const double* locs, double* out,
                   int64_t nbatch, int64_t npts, int ndim,
                   const int64_t* grid, const float* width, int kernel,
                   int norm, int pad, int is_complex, double beta,
                   int nthreads) {
  geom_t G;
  make_geom(&G, ndim, grid, width, kernel, norm, pad);
  const int comp = is_complex ? 2 : 1;
  const int64_t nc = ncells_of(&G);
  (void)nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * G.ntaps_total);
    double* w = (double*)malloc(sizeof(double) * G.ntaps_total);
#pragma omp for schedule(static)
    for (int64_t p = 0; p < npts; ++p) {
      taps_f64(&G, locs + p * ndim, width, beta, cell, w);
      for (int64_t b = 0; b < nbatch; ++b) {
        for (int c = 0; c < comp; ++c) {
          double acc = 0.0;
          for (int64_t j = 0; j < G.ntaps_total; ++j) {
            if (cell[j] < 0) continue;
            acc += w[j] * vals[(b * nc + cell[j]) * comp + c];
          }
          out[(b * npts + p) * comp + c] = acc;
        }
      }
    }
    free(cell);
    free(w);
  }
  return 0;
}

/* --------------------------------------------------------------- grid
 * Sequential scatter-add, batch -> point -> tap, into a zeroed output. */
int orc_grid_f32(const float* vals, const float* locs, float* out,
                 int64_t nbatch, int64_t npts, int ndim, const int64_t* grid,
                 const float* width, int kernel, int norm, int pad,
                 int is_complex, float beta, int nthreads) {
  geom_t G;
  make_geom(&G, ndim, grid, width, kernel, norm, pad);
  const int comp = is_complex ? 2 : 1;
  const int64_t nc = ncells_of(&G);
  memset(out, 0, sizeof(float) * nbatch * nc * comp);
  int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * G.ntaps_total * npts);
  float* w = (float*)malloc(sizeof(float) * G.ntaps_total * npts);
  for (int64_t p = 0; p < npts; ++p)
    taps_f32(&G, locs + p * ndim, width, beta, cell + p * G.ntaps_total,
             w + p * G.ntaps_total);
  /* batches write disjoint outputs: one thread per batch keeps every
   * output's accumulation order identical to the sequential loop */
#pragma omp parallel for schedule(static, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t b = 0; b < nbatch; ++b)
    for (int64_t p = 0; p < npts; ++p)
      for (int64_t j = 0; j < G.ntaps_total; ++j) {
        const int64_t cj = cell[p * G.ntaps_total + j];
        if (cj < 0) continue;
        const float wj = w[p * G.ntaps_total + j];
        for (int c = 0; c < comp; ++c)
          out[(b * nc + cj) * comp + c] =
              out[(b * nc + cj) * comp + c] +
              wj * vals[(b * npts + p) * comp + c];
      }
  free(cell);
  free(w);
  return 0;
}

int orc_grid_f64(const double* vals, const double* locs, double* out,
                 int64_t nbatch, int64_t npts, int ndim, const int64_t* grid,
                 const float* width, int kernel, int norm, int pad,
                 int is_complex, double beta) {
  geom_t G;
  make_geom(&G, ndim, grid, width, kernel, norm, pad);
  const int comp = is_complex ? 2 : 1;
  const int64_t nc = ncells_of(&G);
  memset(out, 0, sizeof(double) * nbatch * nc * comp);
  int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * G.ntaps_total);
  double* w = (double*)malloc(sizeof(double) * G.ntaps_total);
  for (int64_t p = 0; p < npts; ++p) {
    taps_f64(&G, locs + p * ndim, width, beta, cell, w);
    for (int64_t b = 0; b < nbatch; ++b)
      for (int64_t j = 0; j < G.ntaps_total; ++j) {
        if (cell[j] < 0) continue;
        for (int c = 0; c < comp; ++c)
          out[(b * nc + cell[j]) * comp + c] +=
              w[j] * vals[(b * npts + p) * comp + c];
      }
  }
  free(cell);
  free(w);
  return 0;
}

/* ------------------------------------------------------ unfold / fold
 * Sizes are per axis after the reference's complex interleave.  Blocks are
 * visited in lexicographic order; unfold copies, fold accumulates. */
typedef struct {
  int ndim;
  int64_t nb[3], B[3], im[3], s[3];
  int64_t nblk, bsz, imsz;
} bgeom_t;

static void make_bgeom(bgeom_t* g, int ndim, const int64_t* nb,
                       const int64_t* B, const int64_t* im,
                       const int64_t* s) {
  g->ndim = ndim;
  g->nblk = g->bsz = g->imsz = 1;
  for (int d = 0; d < ndim; ++d) {
    g->nb[d] = nb[d];
    g->B[d] = B[d];
    g->im[d] = im[d];
    g->s[d] = s[d];
    g->nblk *= nb[d];
    g->bsz *= B[d];
    g->imsz *= im[d];
  }
}

/* flat image offset of element j (in-block) of block k */
static int64_t img_off(const bgeom_t* g, int64_t k, int64_t j) {
  int64_t bi[3], ji[3];
  for (int d = g->ndim - 1; d >= 0; --d) {
    bi[d] = k % g->nb[d];
    k /= g->nb[d];
    ji[d] = j % g->B[d];
    j /= g->B[d];
  }
  int64_t off = 0;
  for (int d = 0; d < g->ndim; ++d)
    off = off * g->im[d] + bi[d] * g->s[d] + ji[d];
  return off;
}

void orc_unfold(const void* x, void* y, int elem_bytes, int64_t nbatch,
                int ndim, const int64_t* nb, const int64_t* B,
                const int64_t* im, const int64_t* s) {
  bgeom_t g;
  make_bgeom(&g, ndim, nb, B, im, s);
  const char* xc = (const char*)x;
  char* yc = (char*)y;
  for (int64_t n = 0; n < nbatch; ++n)
    for (int64_t k = 0; k < g.nblk; ++k)
      for (int64_t j = 0; j < g.bsz; ++j)
        memcpy(yc + ((n * g.nblk + k) * g.bsz + j) * elem_bytes,
               xc + (n * g.imsz + img_off(&g, k, j)) * elem_bytes,
               elem_bytes);
}

void orc_fold_f32(const float* x, float* y, int64_t nbatch, int ndim,
                  const int64_t* nb, const int64_t* B, const int64_t* im,
                  const int64_t* s) {
  bgeom_t g;
  make_bgeom(&g, ndim, nb, B, im, s);
  memset(y, 0, sizeof(float) * nbatch * g.imsz);
  for (int64_t n = 0; n < nbatch; ++n)
    for (int64_t k = 0; k < g.nblk; ++k)
      for (int64_t j = 0; j < g.bsz; ++j) {
        const int64_t o = n * g.imsz + img_off(&g, k, j);
        y[o] = y[o] + x[(n * g.nblk + k) * g.bsz + j];
      }
}

void orc_fold_f64(const double* x, double* y, int64_t nbatch, int ndim,
                  const int64_t* nb, const int64_t* B, const int64_t* im,
                  const int64_t* s) {
  bgeom_t g;
  make_bgeom(&g, ndim, nb, B, im, s);
  memset(y, 0, sizeof(double) * nbatch * g.imsz);
  for (int64_t n = 0; n < nbatch; ++n)
    for (int64_t k = 0; k < g.nblk; ++k)
      for (int64_t j = 0; j < g.bsz; ++j) {
        const int64_t o = n * g.imsz + img_off(&g, k, j);
        y[o] = y[o] + x[(n * g.nblk + k) * g.bsz + j];
      }
}

/* ------------------------------------------------- trajectory gradient
 * grad[p, d] = sum_b sum_taps Re(conj(h[b, p]) x[b, cell]) dW/dt_{p,d},
 * W the tap weight of taps_f64: norm 1 prod_d win(dn_d), norm 2
 * win(|dn|), dn_d = (t_d - g_d) / (w_d / 2).  win is even in dn, so
 * d win(dn_d)/dt_d = win'(|dn_d|) sign(dn_d) / (w_d / 2). */
static double dwin_f64(int kernel, double u, double beta) {
  if (kernel == KB) return orc_kb_deriv_f64(u, beta);
  return u <= 1.0 ? -1.0 : 0.0;
}

int orc_grad_locs_f64(const double* x, const double* h, const double* locs,
                      double* grad, int64_t nbatch, int64_t npts, int ndim,
                      const int64_t* grid, const float* width, int kernel,
                      int norm, int pad, int is_complex, double beta,
                      int nthreads) {
  geom_t G;
  make_geom(&G, ndim, grid, width, kernel, norm, pad);
  const int comp = is_complex ? 2 : 1;
  const int64_t nc = ncells_of(&G);
  (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t p = 0; p < npts; ++p) {
    const double* t = locs + p * ndim;
    double r[3];
    int64_t lower[3];
    for (int d = 0; d < ndim; ++d) {
      r[d] = (double)width[d] / 2.0;
      lower[d] = (int64_t)ceil(t[d] - r[d]);
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < G.ntaps_total; ++j) {
      int64_t rem = j, k[3];
      for (int d = ndim - 1; d >= 0; --d) {
        k[d] = rem % G.ntap[d];
        rem /= G.ntap[d];
      }
      int keep = 1;
      double dn[3], s2 = 0.0;
      int64_t c = 0;
      for (int d = 0; d < ndim; ++d) {
        int64_t g = lower[d] + k[d];
        double diff = t[d] - (double)g;
        dn[d] = diff / r[d];
        if (!(fabs(diff) <= r[d])) keep = 0;
        if (G.pad == ZERO && (g < 0 || g >= G.n[d])) keep = 0;
        s2 += dn[d] * dn[d];
        c = c * G.n[d] + mod_pos(g, G.n[d]);
      }
      if (!keep) continue;
      double dW[3] = {0.0, 0.0, 0.0};
      if (norm == 1) {
        for (int d = 0; d < ndim; ++d) {
          double prod = 1.0;
          for (int e = 0; e < ndim; ++e)
            if (e != d) prod *= win_f64(kernel, dn[e], beta);
          double sg = dn[d] > 0 ? 1.0 : (dn[d] < 0 ? -1.0 : 0.0);
          dW[d] = dwin_f64(kernel, fabs(dn[d]), beta) * sg / r[d] * prod;
        }
      } else {
        double rr = sqrt(s2);
        if (rr <= 1.0 && rr > 0.0)
          for (int d = 0; d < ndim; ++d)
            dW[d] = dwin_f64(kernel, rr, beta) * dn[d] / (rr * r[d]);
      }
      double s = 0.0;
      for (int64_t b = 0; b < nbatch; ++b)
        for (int cc = 0; cc < comp; ++cc)
          s += h[(b * npts + p) * comp + cc] * x[(b * nc + c) * comp + cc];
      for (int d = 0; d < ndim; ++d) acc[d] += s * dW[d];
    }
    for (int d = 0; d < ndim; ++d) grad[p * ndim + d] = acc[d];
  }
  return 0;
}
