/*
 * pdos_oracle.c -- FP64 CPU restatement of the reference PDOS ADMM iteration.
 *
 * TEST INFRASTRUCTURE.  This library is the parity checker and the timed CPU
 * baseline.  Only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark load it; the product (pdos_amd/, libpdos*_hip.so) never links,
 * loads or calls it.
 *
 * It restates, operation for operation, the reference C code built with
 * gcc -O3 -std=gnu99 -DDLONG -DLDL_LONG (pdos.mk:2-3): sequential sums in
 * storage order, no FMA contraction (this file is compiled for baseline
 * x86-64, which has no FMA), same setup, same stop test.  With the same
 * fill-reducing permutation it reproduces the reference's x, s, y bit for bit
 * (pinned against the compiled reference in tests/golden/).
 *
 * Map to the reference (file:line):
 *   o_setup            common.h:35-218        workspace, NORMALIZE, lambda, x1e3, transpose
 *                      ruiz_equil_common.h:98-322 instead, after oracle_set_equil(1)
 *   o_ruiz_collapse    ruiz_equil_common.h:17-81   cone-collapsed |A|, |b|
 *   o_ruiz             ruiz_equil_common.h:130-269 Ruiz-type scaling of [A b; c' 0]
 *   o_transpose        cs.c:146-168 (+83-97)  stable counting-sort transpose
 *   o_rowsum           linAlg.h:70-115        sequential row sums in CSC/CSR order
 *   o_pri_resid        linAlg.h:158-176
 *   o_dual_resid       linAlg.h:178-191
 *   o_cg               indirect/private.c:31-104
 *   o_lin_indirect     indirect/private.c:20-29,106-118
 *   o_form_kkt         direct/private.c:77-120 (+ cs.c:33-51 compress)
 *   o_symperm          cs.c:109-144
 *   o_ldl_symbolic     direct/ldl.c:187-240   (LDL 2.1, up-looking)
 *   o_ldl_numeric      direct/ldl.c:253-319
 *   o_ldl_solve        direct/ldl.c:326-428 via direct/private.c:200-239
 *   o_lin_direct       direct/private.c:16-56
 *   oracle_proj_cone   cones.c:5-57
 *   oracle_pdos        pdos.c:50-132, 232-303 (loop, relax, cones, dual update, unscale)
 *
 * The only free choice is the fill-reducing ordering of the direct plugin
 * (the reference uses SuiteSparse AMD 2.3, direct/private.c:162-169).  Callers
 * may pass the permutation to use; otherwise o_min_degree (exact minimum
 * degree on a bitset elimination graph) is used.  Any ordering gives the same
 * solution up to roundoff.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/pdos.h"

#define O_SQRT_RATIO 1e3
#define O_RANGE_MAX 1e6
#define O_RANGE_MIN 1e-6

typedef struct {
  long iters;       /* ADMM updates performed */
  long status;      /* 0 Solved, 1 Indeterminate */
  long cg_iters;    /* total CG iterations (indirect) */
  long lnz;         /* nnz(L) (direct) */
  double lambda;
  double p_res, d_res, p_obj, d_obj, eta;  /* last stop check */
  double *hist;     /* optional: 4 doubles per stop check (p_res, d_res, p_obj, d_obj) */
  long hist_cap;    /* capacity of hist in checks */
  long hist_len;    /* checks written */
  double setup_s, solve_s;  /* CLOCK_MONOTONIC seconds: setup(+factor), loop */
} OracleInfo;

typedef struct {
  long n, m;
  /* A (scaled) in CSC: columns of A = rows of A' */
  double *Ax; const long *Ai, *Ap;
  /* A' in CSC = A in CSR: rows of A */
  double *Atx; long *Ati, *Atp;
  double *b, *c;
  double *xbuf;      /* n + max(m, n): x then stilde */
  double *x, *st, *s, *y;
  double *D, *E;
  double lambda;
  const Params *par;
  /* indirect */
  double *cg_p, *cg_q, *cg_Ax;
  /* direct */
  long N;
  long *Lp, *Li; double *Lx, *Ld; long *P; double *bp;
} OWork;

/* Scaling used by o_setup: 0 = common.h (the reference build), 1 = the
 * reference's alternative header ruiz_equil_common.h, which replaces
 * common.h when compiled in its place. */
static int o_equil = 0;

void oracle_set_equil(int e) { o_equil = e; }

static double o_now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static double o_normsq(const double *v, long len) {
  double acc = 0.0;
  long i;
  for (i = 0; i < len; ++i) acc += v[i] * v[i];
  return acc;
}

static double o_dot(const double *a, const double *b, long len) {
  double acc = 0.0;
  long i;
  for (i = 0; i < len; ++i) acc += a[i] * b[i];
  return acc;
}

/* Stable counting-sort transpose of a CSC matrix (cs.c:146-168): within each
 * output column the entries appear in increasing input-column order. */
static void o_transpose(long rows, long cols, const long *Ap, const long *Ai,
                        const double *Ax, long **Tp, long **Ti, double **Tx) {
  long nz = Ap[cols], r, j, p;
  long *cnt = (long *)calloc((size_t)rows + 1, sizeof(long));
  *Tp = (long *)malloc(sizeof(long) * ((size_t)rows + 1));
  *Ti = (long *)malloc(sizeof(long) * (size_t)(nz > 0 ? nz : 1));
  *Tx = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
  for (p = 0; p < nz; ++p) cnt[Ai[p]]++;
  {
    long run = 0;
    for (r = 0; r < rows; ++r) {
      long c = cnt[r];
      (*Tp)[r] = run;
      cnt[r] = run;
      run += c;
    }
    (*Tp)[rows] = run;
  }
  for (j = 0; j < cols; ++j)
    for (p = Ap[j]; p < Ap[j + 1]; ++p) {
      long q = cnt[Ai[p]]++;
      (*Ti)[q] = j;
      (*Tx)[q] = Ax[p];
    }
  free(cnt);
}

/* y[j] (+)= sum_p val[p] * x[idx[p]] over compressed vector j, in storage
 * order (linAlg.h:70-115).  sign = -1 gives the decum variants: each term is
 * negated exactly, so the sum equals the reference's sc = -1.0 product. */
static void o_spmv(long nvec, const long *ptr, const long *idx, const double *val,
                   const double *x, double *y, int sign, int accum) {
  long j, p;
  for (j = 0; j < nvec; ++j) {
    double acc = 0.0;
    if (sign > 0) {
      for (p = ptr[j]; p < ptr[j + 1]; ++p) acc += val[p] * x[idx[p]];
    } else {
      for (p = ptr[j]; p < ptr[j + 1]; ++p) acc += (-val[p]) * x[idx[p]];
    }
    y[j] = accum ? y[j] + acc : acc;
  }
}
/* A*x uses A' in CSC (rows of A); A'*x uses A in CSC (linAlg.h:118-155). */
#define O_ACCUM_A(w, xin, yout, sg, ac) \
  o_spmv((w)->m, (w)->Atp, (w)->Ati, (w)->Atx, (xin), (yout), (sg), (ac))
#define O_ACCUM_AT(w, xin, yout, sg, ac) \
  o_spmv((w)->n, (w)->Ap, (w)->Ai, (w)->Ax, (xin), (yout), (sg), (ac))

/* ruiz_equil_common.h:17-81: in a copy of A's values (ax) and row indices
 * (ai) and of b, every entry of a column that falls in a second-order cone
 * becomes the largest |value| of that column's entries in the cone, with the
 * cone's first row as its row; b's cone entries become the cone's largest
 * |b|.  As in the header, a column's scan for the entries of a cone is bounded
 * by nnz rather than by the column's end, so when the next column begins
 * inside the same cone its leading entries are rewritten too (and its own
 * scan then sees those rewritten values).  A column whose rows are not sorted
 * stops at the last cone (the header reads past q[] there). */
static void o_ruiz_collapse(const Data *d, const Cone *k, double *ax, long *ai, double *b) {
  const long n = d->n, nz = d->Ap[n], top = k->f + k->l;
  for (long col = 0; col < n; ++col) {
    long r0 = top, p = d->Ap[col], cone = 0;
    while (p < d->Ap[col + 1] && d->Ai[p] < r0) ++p;
    while (p < d->Ap[col + 1] && cone < k->qsize) {
      const long r1 = r0 + k->q[cone];
      double mx = 0.0;
#ifdef ORACLE_MUTANT_RUIZ_SCAN
      const long lim = d->Ap[col + 1];  /* planted: the scan stops at the column's end */
#else
      const long lim = nz;
#endif
      for (long t = p; t < lim && d->Ai[t] >= r0 && d->Ai[t] < r1; ++t) {
        const double v = fabs(ax[t]);
        mx = mx > v ? mx : v;
      }
      for (; p < lim && d->Ai[p] >= r0 && d->Ai[p] < r1; ++p) {
        ax[p] = mx;
        ai[p] = r0;
      }
      r0 = r1;
      ++cone;
    }
  }
  for (long cone = 0, r0 = top; cone < k->qsize; r0 += k->q[cone], ++cone) {
    double mx = 0.0;
    for (long i = 0; i < k->q[cone]; ++i) {
      const double v = fabs(b[r0 + i]);
      mx = mx > v ? mx : v;
    }
    for (long i = 0; i < k->q[cone]; ++i) b[r0 + i] = mx;
  }
}

/* ruiz_equil_common.h:130-269: three sweeps over [A b; c' 0] (A, b
 * collapsed as above).  Each sweep takes the row maxima pi_i = max(max_j
 * |A_ij| E_j, |b_i| e) with the last row's pi = max_j |c_j| E_j, gives every
 * row of a cone its first row's pi, and sets D_i = sqrt(D_i / pi_i) (1 where
 * pi_i <= 1e-6); then the column maxima delta_j = max(max_i |A_ij| D_i, |c_j| d)
 * with the last column's delta = max_i |b_i| D_i, and E_j = sqrt(E_j /
 * delta_j) (1 where delta_j <= 1e-6); d and e are the last row's D and last
 * column's E.  Then A, b, c and the warm start are scaled with D and E. */
static void o_ruiz(const Data *d, const Cone *k, OWork *w) {
  const long n = d->n, m = d->m, nz = d->Ap[n];
  double *ax = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
  long *ai = (long *)malloc(sizeof(long) * (size_t)(nz > 0 ? nz : 1));
  double *bc = (double *)malloc(sizeof(double) * (size_t)m);
  double *pi = (double *)calloc((size_t)m, sizeof(double));
  double *delta = (double *)calloc((size_t)n, sizeof(double));
  memcpy(ax, d->Ax, sizeof(double) * (size_t)nz);
  memcpy(ai, d->Ai, sizeof(long) * (size_t)nz);
  memcpy(bc, d->b, sizeof(double) * (size_t)m);
  if (k->qsize > 0) o_ruiz_collapse(d, k, ax, ai, bc);
  for (long i = 0; i < m; ++i) w->D[i] = 1.0;
  for (long j = 0; j < n; ++j) w->E[j] = 1.0;
  double dl = 1.0, el = 1.0, pil = 0.0, dlt = 0.0;
  for (int sweep = 0; sweep < 3; ++sweep) {
    for (long j = 0; j < n; ++j) {
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) {
        const double v = fabs(ax[p]) * w->E[j];
        pi[ai[p]] = pi[ai[p]] > v ? pi[ai[p]] : v;
      }
      const double v = fabs(d->c[j]) * w->E[j];
      pil = pil > v ? pil : v;
    }
    for (long i = 0; i < m; ++i) {
      const double v = fabs(bc[i]) * el;
      pi[i] = pi[i] > v ? pi[i] : v;
    }
#ifndef ORACLE_MUTANT_RUIZ_CONE  /* planted: cone rows keep their own maxima */
    for (long cone = 0, r0 = k->f + k->l; cone < k->qsize; r0 += k->q[cone], ++cone)
      for (long i = 0; i < k->q[cone]; ++i) pi[r0 + i] = pi[r0];
#endif
    for (long i = 0; i < m; ++i) {
      w->D[i] = fabs(pi[i]) > 1e-6 ? sqrt(w->D[i] / pi[i]) : 1.0;
      pi[i] = 0.0;
    }
    dl = fabs(pil) > 1e-6 ? sqrt(dl / pil) : 1.0;
    pil = 0.0;
    for (long j = 0; j < n; ++j) {
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) {
        const double v = fabs(ax[p]) * w->D[ai[p]];
        delta[j] = delta[j] > v ? delta[j] : v;
      }
      const double v = fabs(d->c[j]) * dl;
      delta[j] = delta[j] > v ? delta[j] : v;
    }
    for (long i = 0; i < m; ++i) {
      const double v = fabs(bc[i]) * w->D[i];
      dlt = dlt > v ? dlt : v;
    }
    for (long j = 0; j < n; ++j) {
      w->E[j] = fabs(delta[j]) > 1e-6 ? sqrt(w->E[j] / delta[j]) : 1.0;
      delta[j] = 0.0;
    }
    el = fabs(dlt) > 1e-6 ? sqrt(el / dlt) : 1.0;
    dlt = 0.0;
  }
  for (long j = 0; j < n; ++j)
    for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) w->Ax[p] = w->D[d->Ai[p]] * d->Ax[p] * w->E[j];
  for (long j = 0; j < n; ++j) w->c[j] = d->c[j] * w->E[j];
  for (long i = 0; i < m; ++i) w->b[i] = d->b[i] * w->D[i];
  for (long i = 0; i < m; ++i) w->y[i] /= w->D[i];
  for (long j = 0; j < n; ++j) w->x[j] /= w->E[j];
  for (long i = 0; i < m; ++i) w->s[i] *= w->D[i];
  free(ax); free(ai); free(bc); free(pi); free(delta);
}

/* Workspace setup, common.h:35-218.  The caller's arrays are copied, never
 * written (the reference scales d->c / d->Ax in place when NORMALIZE == 0;
 * the values it computes with are identical to these copies). */
static OWork *o_setup(const Data *d, const Cone *k) {
  long n = d->n, m = d->m, nz = d->Ap[d->n], i, j;
  OWork *w = (OWork *)calloc(1, sizeof(OWork));
  w->n = n; w->m = m; w->par = d->p;
  w->xbuf = (double *)calloc((size_t)(n + (m > n ? m : n)), sizeof(double));
  w->x = w->xbuf; w->st = w->xbuf + n;
  if (d->x) memcpy(w->x, d->x, sizeof(double) * (size_t)n);
  w->s = (double *)calloc((size_t)m + 1, sizeof(double));
  if (d->s) memcpy(w->s, d->s, sizeof(double) * (size_t)m);
  w->y = (double *)calloc((size_t)m + 1, sizeof(double));
  if (d->y) memcpy(w->y, d->y, sizeof(double) * (size_t)m);
  w->D = (double *)calloc((size_t)m + 1, sizeof(double));
  w->E = (double *)calloc((size_t)n + 1, sizeof(double));
  w->Ax = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
  memcpy(w->Ax, d->Ax, sizeof(double) * (size_t)nz);
  w->Ai = d->Ai; w->Ap = d->Ap;
  w->b = (double *)malloc(sizeof(double) * (size_t)(m + 1));
  w->c = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  memcpy(w->b, d->b, sizeof(double) * (size_t)m);
  memcpy(w->c, d->c, sizeof(double) * (size_t)n);

  if (o_equil == 1 && d->p->NORMALIZE) {
    o_ruiz(d, k, w);
  } else if (d->p->NORMALIZE) {
    long row0 = k->f + k->l, cone;
    /* squared row norms of A (common.h:89-93) */
    for (j = 0; j < n; ++j)
      for (i = d->Ap[j]; i < d->Ap[j + 1]; ++i) w->D[d->Ai[i]] += d->Ax[i] * d->Ax[i];
    /* one value per SOC: the block mean (common.h:95-105) */
    for (cone = 0; cone < k->qsize; ++cone) {
      double acc = 0.0;
      long q = k->q[cone];
      for (i = 0; i < q; ++i) acc += w->D[row0 + i];
      for (i = 0; i < q; ++i) w->D[row0 + i] = acc / q;
      row0 += q;
    }
    for (i = 0; i < m; ++i) {
      double v = fabs(w->D[i]);
      v = v > O_RANGE_MIN ? v : O_RANGE_MIN;
      v = v > O_RANGE_MAX ? O_RANGE_MAX : v;
      w->D[i] = 1.0 / sqrt(v);
    }
    for (j = 0; j < n; ++j)
      for (i = d->Ap[j]; i < d->Ap[j + 1]; ++i) w->Ax[i] *= w->D[d->Ai[i]];
    for (j = 0; j < n; ++j)
      for (i = d->Ap[j]; i < d->Ap[j + 1]; ++i) w->E[j] += w->Ax[i] * w->Ax[i];
    for (j = 0; j < n; ++j) {
      double v = fabs(w->E[j]);
      v = v > O_RANGE_MIN ? v : O_RANGE_MIN;
      v = v > O_RANGE_MAX ? O_RANGE_MAX : v;
      w->E[j] = 1.0 / sqrt(v);
    }
    for (j = 0; j < n; ++j)
      for (i = d->Ap[j]; i < d->Ap[j + 1]; ++i) w->Ax[i] *= w->E[j];
    for (j = 0; j < n; ++j) w->c[j] = d->c[j] * w->E[j];
    for (i = 0; i < m; ++i) w->b[i] = d->b[i] * w->D[i];
    for (i = 0; i < m; ++i) w->y[i] /= w->D[i];
    for (j = 0; j < n; ++j) w->x[j] /= w->E[j];
    for (i = 0; i < m; ++i) w->s[i] *= w->D[i];
  } else {
    for (i = 0; i < m; ++i) w->D[i] = 1.0;
    for (j = 0; j < n; ++j) w->E[j] = 1.0;
  }
  {
    const double eps = o_equil == 1 ? 1e-6 : 1e-15;  /* ruiz_equil_common.h:288, common.h:182 */
    w->lambda = (eps + sqrt(o_normsq(w->b, m))) / (eps + sqrt(o_normsq(w->c, n)));
  }
  for (j = 0; j < n; ++j) {
    w->E[j] *= O_SQRT_RATIO;
    w->c[j] *= O_SQRT_RATIO;
    w->x[j] /= O_SQRT_RATIO;
  }
  for (i = 0; i < nz; ++i) w->Ax[i] *= O_SQRT_RATIO;
  o_transpose(m, n, w->Ap, w->Ai, w->Ax, &w->Atp, &w->Ati, &w->Atx);
  return w;
}

/* ||(s - b + A x) ./ D||_2, using stilde as scratch (linAlg.h:158-176). */
static double o_pri_resid(OWork *w) {
  long i;
  for (i = 0; i < w->m; ++i) w->st[i] = w->s[i] - w->b[i];
  O_ACCUM_A(w, w->x, w->st, +1, 1);
#ifdef ORACLE_MUTANT_RES_D
  for (i = 0; i < w->m; ++i) w->st[i] *= w->D[i];
#else
  for (i = 0; i < w->m; ++i) w->st[i] /= w->D[i];
#endif
  return sqrt(o_normsq(w->st, w->m));
}

/* ||(c + A' y) ./ E||_2 (linAlg.h:178-191). */
static double o_dual_resid(OWork *w) {
  long i;
  memcpy(w->st, w->c, sizeof(double) * (size_t)w->n);
  O_ACCUM_AT(w, w->y, w->st, +1, 1);
  for (i = 0; i < w->n; ++i) w->st[i] /= w->E[i];
  return sqrt(o_normsq(w->st, w->n));
}

/* ---------------- indirect plugin: CG on (I + A'A) ---------------- */

/* indirect/private.c:31-104. Returns the CG iterations performed. */
static long o_cg(OWork *w) {
  const long maxit = w->par->CG_MAX_ITS, n = w->n;
  const double tol_sq = w->par->CG_TOL * w->par->CG_TOL;
  double *p = w->cg_p, *q = w->cg_q, *Ax = w->cg_Ax, *x = w->x;
  double rho, rho_new = 0.0, alpha, beta;
  long i = 0, j;
  memcpy(Ax, w->st, sizeof(double) * (size_t)w->m);
  for (j = 0; j < n; ++j) q[j] = (-w->lambda) * w->c[j];
  O_ACCUM_A(w, x, Ax, +1, 1);
#ifdef ORACLE_MUTANT_CG_SIGN
  O_ACCUM_AT(w, Ax, q, +1, 1);
#else
  O_ACCUM_AT(w, Ax, q, -1, 1);
#endif
  memcpy(p, q, sizeof(double) * (size_t)n);
  rho = o_normsq(q, n);
  if (rho > tol_sq) {
    for (i = 0; i < maxit; ++i) {
      O_ACCUM_A(w, p, Ax, +1, 0);
      memcpy(w->st, p, sizeof(double) * (size_t)n);
      O_ACCUM_AT(w, Ax, w->st, +1, 1);
      alpha = rho / o_dot(p, w->st, n);
      for (j = 0; j < n; ++j) x[j] += alpha * p[j];
      for (j = 0; j < n; ++j) q[j] += (-alpha) * w->st[j];
      rho_new = o_normsq(q, n);
      if (rho_new < tol_sq) {
        ++i;
        break;
      }
      beta = rho_new / rho;
      for (j = 0; j < n; ++j) p[j] *= beta;
      for (j = 0; j < n; ++j) p[j] += 1.0 * q[j];
      rho = rho_new;
    }
  }
  return i;
}

/* indirect/private.c:20-29, 106-118. */
static long o_lin_indirect(OWork *w) {
  long i, its;
  for (i = 0; i < w->m; ++i) w->st[i] = w->s[i] + w->lambda * w->y[i] - w->b[i];
  its = o_cg(w);
  memcpy(w->st, w->b, sizeof(double) * (size_t)w->m);
  O_ACCUM_A(w, w->x, w->st, -1, 1);
  return its;
}

/* ---------------- direct plugin: LDL' of the quasi-definite KKT ---------------- */

/* Upper triangle of K = [[-I, A'], [A, I]] in CSC, entries of each column in
 * triplet order (direct/private.c:77-120, cs.c:33-51). */
static void o_form_kkt(const OWork *w, long **Kp, long **Ki, double **Kx) {
  long n = w->n, m = w->m, N = n + m, nz = w->Ap[n], Knz = n + nz + m;
  long *tr = (long *)malloc(sizeof(long) * (size_t)Knz);
  long *tc = (long *)malloc(sizeof(long) * (size_t)Knz);
  double *tv = (double *)malloc(sizeof(double) * (size_t)Knz);
  long *cnt = (long *)calloc((size_t)N + 1, sizeof(long));
  long t = 0, j, p;
  for (j = 0; j < n; ++j) { tr[t] = j; tc[t] = j; tv[t] = -1.0; ++t; }
  for (j = 0; j < n; ++j)
    for (p = w->Ap[j]; p < w->Ap[j + 1]; ++p) {
      tr[t] = j; tc[t] = w->Ai[p] + n; tv[t] = w->Ax[p]; ++t;
    }
  for (j = 0; j < m; ++j) { tr[t] = j + n; tc[t] = j + n; tv[t] = 1.0; ++t; }
  *Kp = (long *)malloc(sizeof(long) * (size_t)(N + 1));
  *Ki = (long *)malloc(sizeof(long) * (size_t)Knz);
  *Kx = (double *)malloc(sizeof(double) * (size_t)Knz);
  for (t = 0; t < Knz; ++t) cnt[tc[t]]++;
  {
    long run = 0;
    for (j = 0; j < N; ++j) { long c = cnt[j]; (*Kp)[j] = run; cnt[j] = run; run += c; }
    (*Kp)[N] = run;
  }
  for (t = 0; t < Knz; ++t) {
    long q = cnt[tc[t]]++;
    (*Ki)[q] = tr[t];
    (*Kx)[q] = tv[t];
  }
  free(tr); free(tc); free(tv); free(cnt);
}

/* Upper triangle of P K P' (cs.c:109-144): column max(i2,j2) receives row
 * min(i2,j2), entries in (column, position) visiting order of K. */
static void o_symperm(long N, const long *Kp, const long *Ki, const double *Kx,
                      const long *pinv, long **Cp, long **Ci, double **Cx) {
  long *cnt = (long *)calloc((size_t)N + 1, sizeof(long));
  long nz = Kp[N], j, p;
  *Cp = (long *)malloc(sizeof(long) * (size_t)(N + 1));
  *Ci = (long *)malloc(sizeof(long) * (size_t)(nz > 0 ? nz : 1));
  *Cx = (double *)malloc(sizeof(double) * (size_t)(nz > 0 ? nz : 1));
  for (j = 0; j < N; ++j) {
    long j2 = pinv[j];
    for (p = Kp[j]; p < Kp[j + 1]; ++p) {
      long i = Ki[p], i2;
      if (i > j) continue;
      i2 = pinv[i];
      cnt[i2 > j2 ? i2 : j2]++;
    }
  }
  {
    long run = 0;
    for (j = 0; j < N; ++j) { long c = cnt[j]; (*Cp)[j] = run; cnt[j] = run; run += c; }
    (*Cp)[N] = run;
  }
  for (j = 0; j < N; ++j) {
    long j2 = pinv[j];
    for (p = Kp[j]; p < Kp[j + 1]; ++p) {
      long i = Ki[p], i2, q;
      if (i > j) continue;
      i2 = pinv[i];
      q = cnt[i2 > j2 ? i2 : j2]++;
      (*Ci)[q] = i2 < j2 ? i2 : j2;
      (*Cx)[q] = Kx[p];
    }
  }
  free(cnt);
}

/* Elimination tree and column counts of L for an upper-triangular CSC
 * pattern (the up-looking symbolic pass, ldl.c:187-240, with no permutation). */
static void o_ldl_symbolic(long N, const long *Cp, const long *Ci, long *Lp, long *parent,
                           long *lnz, long *flag) {
  long k, p;
  for (k = 0; k < N; ++k) {
    parent[k] = -1;
    flag[k] = k;
    lnz[k] = 0;
    for (p = Cp[k]; p < Cp[k + 1]; ++p) {
      long i = Ci[p];
      if (i < k) {
        while (flag[i] != k) {
          if (parent[i] == -1) parent[i] = k;
          lnz[i]++;
          flag[i] = k;
          i = parent[i];
        }
      }
    }
  }
  Lp[0] = 0;
  for (k = 0; k < N; ++k) Lp[k + 1] = Lp[k] + lnz[k];
}

/* Numeric up-looking LDL' (ldl.c:253-319): row k of L is a sparse triangular
 * solve over the etree reach of column k, visited in topological order. */
static long o_ldl_numeric(long N, const long *Cp, const long *Ci, const double *Cx,
                          const long *Lp, const long *parent, long *lnz, long *Li,
                          double *Lx, double *Dv, double *Y, long *pattern, long *flag) {
  long k, p;
  for (k = 0; k < N; ++k) {
    long top = N;
    Y[k] = 0.0;
    flag[k] = k;
    lnz[k] = 0;
    for (p = Cp[k]; p < Cp[k + 1]; ++p) {
      long i = Ci[p], len = 0;
      if (i <= k) {
        Y[i] += Cx[p];
        while (flag[i] != k) {
          pattern[len++] = i;
          flag[i] = k;
          i = parent[i];
        }
        while (len > 0) pattern[--top] = pattern[--len];
      }
    }
    Dv[k] = Y[k];
    Y[k] = 0.0;
    for (; top < N; ++top) {
      long i = pattern[top], pend, q;
      double yi = Y[i], lki;
      Y[i] = 0.0;
      pend = Lp[i] + lnz[i];
      for (q = Lp[i]; q < pend; ++q) Y[Li[q]] -= Lx[q] * yi;
      lki = yi / Dv[i];
      Dv[k] -= lki * yi;
      Li[q] = k;
      Lx[q] = lki;
      lnz[i]++;
    }
    if (Dv[k] == 0.0) return k;
  }
  return N;
}

/* Exact minimum degree on a bitset elimination graph (ties -> lowest index).
 * perm[k] = original index of the k-th pivot.  Returns 0, or -1 if N is too
 * large for the dense bitset (the oracle refuses rather than thrash). */
static int o_min_degree(long N, const long *Kp, const long *Ki, long *perm) {
  long W = (N + 63) / 64, j, p, step;
  uint64_t *adj;
  long *deg;
  char *dead;
  if (N > 40000) return -1;
  adj = (uint64_t *)calloc((size_t)(N * W), sizeof(uint64_t));
  deg = (long *)calloc((size_t)N, sizeof(long));
  dead = (char *)calloc((size_t)N, 1);
  for (j = 0; j < N; ++j)
    for (p = Kp[j]; p < Kp[j + 1]; ++p) {
      long i = Ki[p];
      if (i == j) continue;
      adj[i * W + j / 64] |= (uint64_t)1 << (j % 64);
      adj[j * W + i / 64] |= (uint64_t)1 << (i % 64);
    }
  for (j = 0; j < N; ++j) {
    long c = 0, t;
    for (t = 0; t < W; ++t) c += __builtin_popcountll(adj[j * W + t]);
    deg[j] = c;
  }
  for (step = 0; step < N; ++step) {
    long v = -1, best = -1, t;
    uint64_t *av;
    for (j = 0; j < N; ++j)
      if (!dead[j] && (v < 0 || deg[j] < best)) { v = j; best = deg[j]; }
    perm[step] = v;
    dead[v] = 1;
    av = adj + v * W;
    for (t = 0; t < W; ++t) {
      uint64_t bits = av[t];
      while (bits) {
        long u = t * 64 + __builtin_ctzll(bits);
        uint64_t *au = adj + u * W;
        long c = 0, s;
        bits &= bits - 1;
        for (s = 0; s < W; ++s) { au[s] |= av[s]; }
        au[u / 64] &= ~((uint64_t)1 << (u % 64));
        au[v / 64] &= ~((uint64_t)1 << (v % 64));
        for (s = 0; s < W; ++s) c += __builtin_popcountll(au[s]);
        deg[u] = c;
      }
    }
  }
  free(adj); free(deg); free(dead);
  return 0;
}

static int o_factor(OWork *w, const long *perm_in) {
  long n = w->n, m = w->m, N = n + m, k;
  long *Kp, *Ki, *Cp, *Ci, *pinv, *parent, *lnz, *flag, *pattern;
  double *Kx, *Cx, *Y;
  w->N = N;
  o_form_kkt(w, &Kp, &Ki, &Kx);
  w->P = (long *)malloc(sizeof(long) * (size_t)N);
  if (perm_in) {
    memcpy(w->P, perm_in, sizeof(long) * (size_t)N);
  } else if (o_min_degree(N, Kp, Ki, w->P) != 0) {
    free(Kp); free(Ki); free(Kx);
    return -1;
  }
  pinv = (long *)malloc(sizeof(long) * (size_t)N);
  for (k = 0; k < N; ++k) pinv[w->P[k]] = k;
  o_symperm(N, Kp, Ki, Kx, pinv, &Cp, &Ci, &Cx);
  parent = (long *)malloc(sizeof(long) * (size_t)N);
  lnz = (long *)malloc(sizeof(long) * (size_t)N);
  flag = (long *)malloc(sizeof(long) * (size_t)N);
  pattern = (long *)malloc(sizeof(long) * (size_t)N);
  Y = (double *)malloc(sizeof(double) * (size_t)N);
  w->Lp = (long *)malloc(sizeof(long) * (size_t)(N + 1));
  o_ldl_symbolic(N, Cp, Ci, w->Lp, parent, lnz, flag);
  w->Li = (long *)malloc(sizeof(long) * (size_t)(w->Lp[N] > 0 ? w->Lp[N] : 1));
  w->Lx = (double *)malloc(sizeof(double) * (size_t)(w->Lp[N] > 0 ? w->Lp[N] : 1));
  w->Ld = (double *)malloc(sizeof(double) * (size_t)N);
  o_ldl_numeric(N, Cp, Ci, Cx, w->Lp, parent, lnz, w->Li, w->Lx, w->Ld, Y, pattern, flag);
  w->bp = (double *)malloc(sizeof(double) * (size_t)N);
  free(Kp); free(Ki); free(Kx); free(Cp); free(Ci); free(Cx); free(pinv);
  free(parent); free(lnz); free(flag); free(pattern); free(Y);
  return 0;
}

/* x = P' L^-T D^-1 L^-1 P b, in place on the n+m buffer (direct/private.c:200-239). */
static void o_ldl_solve(OWork *w, double *x) {
  long N = w->N, j, p;
  double *bp = w->bp;
  for (j = 0; j < N; ++j) bp[j] = x[w->P[j]];
#ifdef ORACLE_MUTANT_LDL_ORDER
  for (j = N - 1; j >= 0; --j)
#else
  for (j = 0; j < N; ++j)
#endif
    for (p = w->Lp[j]; p < w->Lp[j + 1]; ++p) bp[w->Li[p]] -= w->Lx[p] * bp[j];
  for (j = 0; j < N; ++j) bp[j] /= w->Ld[j];
  for (j = N - 1; j >= 0; --j)
    for (p = w->Lp[j]; p < w->Lp[j + 1]; ++p) bp[j] -= w->Lx[p] * bp[w->Li[p]];
  for (j = 0; j < N; ++j) x[w->P[j]] = bp[j];
}

/* direct/private.c:16-56 */
static void o_lin_direct(OWork *w) {
  long i;
  for (i = 0; i < w->n; ++i) w->x[i] = w->lambda * w->c[i] - w->x[i];
  for (i = 0; i < w->m; ++i) w->st[i] = w->b[i] - (w->s[i] + w->lambda * w->y[i]);
  o_ldl_solve(w, w->xbuf);
  for (i = 0; i < w->m; ++i) w->st[i] += (w->s[i] + w->lambda * w->y[i]);
}

/* ---------------- cone projection, cones.c:5-57 ---------------- */

void oracle_proj_cone(double *x, const Cone *k) {
  long i, j, row;
  memset(x, 0, sizeof(double) * (size_t)k->f);
  for (i = k->f; i < k->f + k->l; ++i)
    if (x[i] < 0.0) x[i] = 0.0;
  row = k->f + k->l;
  for (i = 0; i < k->qsize; ++i) {
    long q = k->q[i];
    double t = x[row], ns = 0.0, a;
    for (j = 0; j < q - 1; ++j) ns += x[row + j + 1] * x[row + j + 1];
    ns = sqrt(ns);
#ifdef ORACLE_MUTANT_SOC_HALF
    a = (ns + t);
#else
    a = (ns + t) / 2.0;
#endif
    if (ns <= t) {
    } else if (ns <= -t) {
      for (j = 0; j < q; ++j) x[row + j] = 0.0;
    } else {
      x[row] = a;
      for (j = 0; j < q - 1; ++j) x[row + j + 1] = a * (x[row + j + 1]) / ns;
    }
    row += q;
  }
}

static void o_free(OWork *w) {
  if (!w) return;
  free(w->xbuf); free(w->s); free(w->y); free(w->D); free(w->E);
  free(w->Ax); free(w->b); free(w->c); free(w->Atx); free(w->Ati); free(w->Atp);
  free(w->cg_p); free(w->cg_q); free(w->cg_Ax);
  free(w->Lp); free(w->Li); free(w->Lx); free(w->Ld); free(w->P); free(w->bp);
  free(w);
}

/* Setup products exposed for pinning common.h: D[m], E[n] (including the
 * x1e3 on E) and lambda. */
int oracle_setup(const Data *d, const Cone *k, double *D, double *E, double *lambda) {
  OWork *w;
  if (!d || !k) return -1;
  w = o_setup(d, k);
  memcpy(D, w->D, sizeof(double) * (size_t)d->m);
  memcpy(E, w->E, sizeof(double) * (size_t)d->n);
  *lambda = w->lambda;
  o_free(w);
  return 0;
}

/* Fill-reducing ordering the oracle uses when none is given (for tests that
 * want to compare factorizations under one ordering). */
int oracle_kkt_order(const Data *d, const Cone *k, long *perm) {
  OWork *w;
  long *Kp, *Ki;
  double *Kx;
  int rc;
  if (!d || !k) return -1;
  w = o_setup(d, k);
  o_form_kkt(w, &Kp, &Ki, &Kx);
  rc = o_min_degree(d->n + d->m, Kp, Ki, perm);
  free(Kp); free(Ki); free(Kx);
  o_free(w);
  return rc;
}

/*
 * The solve (pdos.c:50-132).  plugin 0 = indirect (CG), 1 = direct (LDL').
 * perm: optional KKT permutation for the direct plugin.  info: optional.
 * Returns a malloc'd Sol (release with oracle_free_sol), NULL on bad input.
 */
Sol *oracle_pdos(const Data *d, const Cone *k, int plugin, const long *perm, OracleInfo *info) {
  const Params *par;
  OWork *w;
  Sol *sol;
  long i, j, state = 1, cg_total = 0;
  double pscale, dscale, pres = -1, dres = -1, pobj = -1, dobj = -1, eta = -1;
  double t0, t1, t2;
  if (!d || !k) return NULL;
  {
    long tot = k->f + k->l;
    for (i = 0; i < k->qsize; ++i) tot += k->q[i];
    if (tot != d->m) return NULL;
  }
  par = d->p;
  t0 = o_now();
  pscale = 1.0 + sqrt(o_normsq(d->b, d->m));
  dscale = 1.0 + sqrt(o_normsq(d->c, d->n));
  w = o_setup(d, k);
  if (plugin == 0) {
    w->cg_Ax = (double *)calloc((size_t)d->m + 1, sizeof(double));
    w->cg_p = (double *)calloc((size_t)d->n + 1, sizeof(double));
    w->cg_q = (double *)calloc((size_t)d->n + 1, sizeof(double));
  } else if (o_factor(w, perm) != 0) {
    o_free(w);
    return NULL;
  }
  t1 = o_now();
  if (info) info->hist_len = 0;
  for (i = 0; i < par->MAX_ITERS; ++i) {
    pres = o_pri_resid(w);
    dres = o_dual_resid(w);
    pobj = o_dot(w->c, w->x, w->n);
    dobj = -o_dot(w->b, w->y, w->m);
    eta = fabs(pobj - dobj);
    if (info && info->hist && info->hist_len < info->hist_cap) {
      double *h = info->hist + 4 * info->hist_len;
      h[0] = pres; h[1] = dres; h[2] = pobj; h[3] = dobj;
      info->hist_len++;
    }
    if (pres < par->EPS_ABS * pscale && dres < par->EPS_ABS * dscale &&
        eta < par->EPS_ABS * (1.0 + fabs(pobj) + fabs(dobj))) {
      state = 0;
      break;
    }
    if (plugin == 0) cg_total += o_lin_indirect(w);
    else o_lin_direct(w);
#ifdef ORACLE_MUTANT_RELAX
    for (j = 0; j < w->m; ++j) w->st[j] = par->ALPHA * w->st[j];
#else
    for (j = 0; j < w->m; ++j) w->st[j] = par->ALPHA * w->st[j] + (1.0 - par->ALPHA) * w->s[j];
#endif
    for (j = 0; j < w->m; ++j) w->s[j] = w->st[j] - w->lambda * w->y[j];
    oracle_proj_cone(w->s, k);
#ifdef ORACLE_MUTANT_DUAL_LAMBDA
    for (j = 0; j < w->m; ++j) w->y[j] += (w->s[j] - w->st[j]) * w->lambda;
#else
    for (j = 0; j < w->m; ++j) w->y[j] += (w->s[j] - w->st[j]) / w->lambda;
#endif
  }
  t2 = o_now();
  sol = (Sol *)malloc(sizeof(Sol));
  sol->n = w->n; sol->m = w->m;
  sol->x = (double *)malloc(sizeof(double) * (size_t)(w->n + 1));
  sol->y = (double *)malloc(sizeof(double) * (size_t)(w->m + 1));
  sol->s = (double *)malloc(sizeof(double) * (size_t)(w->m + 1));
  for (j = 0; j < w->n; ++j) sol->x[j] = w->E[j] * w->x[j];
  for (j = 0; j < w->m; ++j) sol->y[j] = w->D[j] * w->y[j];
  for (j = 0; j < w->m; ++j) sol->s[j] = w->s[j] / w->D[j];
  memset(sol->status, 0, sizeof(sol->status));
  strcpy(sol->status, state == 0 ? "Solved" : "Indeterminate");
  if (info) {
    info->iters = i;
    info->status = state;
    info->cg_iters = cg_total;
    info->lnz = plugin == 1 ? w->Lp[w->N] : 0;
    info->lambda = w->lambda;
    info->p_res = pres; info->d_res = dres; info->p_obj = pobj; info->d_obj = dobj;
    info->eta = eta;
    info->setup_s = t1 - t0;
    info->solve_s = t2 - t1;
  }
  o_free(w);
  return sol;
}

void oracle_free_sol(Sol *s) {
  if (!s) return;
  free(s->x); free(s->y); free(s->s); free(s);
}

/* Batch helper for the CPU baseline: solves `count` problems back to back on
 * the calling thread; returns the number that reached "Solved". */
long oracle_pdos_many(const Data *const *ds, const Cone *const *ks, long count, int plugin,
                      long *iters_out) {
  long i, solved = 0;
  for (i = 0; i < count; ++i) {
    OracleInfo info;
    Sol *s;
    memset(&info, 0, sizeof(info));
    s = oracle_pdos(ds[i], ks[i], plugin, NULL, &info);
    if (s) {
      if (info.status == 0) ++solved;
      if (iters_out) iters_out[i] = info.iters;
      oracle_free_sol(s);
    } else if (iters_out) {
      iters_out[i] = -1;
    }
  }
  return solved;
}
