/*
 * pdos.h -- public C ABI of the MI355X PDOS solver.
 *
 * Drop-in for the reference's three-function API.  Struct layouts are
 * ABI-identical to the reference built with -DDLONG (idxint = long, LP64),
 * so a caller compiled against the reference headers links against
 * libpdosindir_hip.so / libpdosdir_hip.so unchanged.
 *
 *   reference interface                     replaced by
 *   pdos.h:15-19   Params (56 B)            Params
 *   pdos.h:22-35   Data   (88 B, x,y,s)     Data
 *   pdos.h:37-41   Sol    (56 B, x,s,y)     Sol
 *   cones.h:8-13   Cone   (32 B)            Cone
 *   globals.h:30-40 idxint (long)           idxint
 *   pdos.h:85      pdos()                   pdos()
 *   pdos.h:86      freeData()               freeData()
 *   pdos.h:87      freeSol()                freeSol()
 *   util.h:37-47   timer, tic, toc, tocq    timer, tic, toc, tocq
 *
 * Semantics (same as the reference unless noted):
 *   - pdos(NULL, k) / pdos(d, NULL) return NULL (pdos.c:61-63).
 *   - m != f + l + sum(q) aborts via assert (pdos.c:65-71).
 *   - Sol and its x[n], s[m], y[m] are malloc'd; release with freeSol().
 *   - freeData() free()s every array of Data/Cone and the structs.
 *   - Deviation: caller data is never written (the reference rescales
 *     d->c and d->Ax in place when NORMALIZE == 0, common.h:169-192, and
 *     restores them with a possibly 1-ulp lossy division, pdos.c:170-177).
 *   - Deviation: reentrant; each call owns its own device arena and stream.
 *   - A GPU failure returns NULL; pdos_last_error() (pdos_ext.h) says why.
 */
#ifndef PDOS_AMD_PDOS_H
#define PDOS_AMD_PDOS_H

#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef long idxint;

typedef struct PROBLEM_PARAMS {
  idxint MAX_ITERS, CG_MAX_ITS;
  double EPS_ABS, ALPHA, CG_TOL;
  idxint VERBOSE, NORMALIZE;
} Params;

typedef struct PROBLEM_DATA {
  idxint n, m;          /* A is m x n */
  double *Ax;           /* A in CSC: values [Ap[n]] */
  idxint *Ai, *Ap;      /* row indices [Ap[n]], column pointers [n+1] */
  double *b, *c;        /* b[m], c[n] */
  double *x;            /* optional warm start x[n]  (may be NULL) */
  double *y;            /* optional warm start y[m]  (may be NULL) */
  double *s;            /* optional warm start s[m]  (may be NULL) */
  Params *p;
} Data;

typedef struct SOL_VARS {
  idxint n, m;
  double *x, *s, *y;
  char status[16];      /* "Solved" or "Indeterminate" */
} Sol;

typedef struct CONE {
  idxint f;             /* zero-cone rows (equalities), first in row order */
  idxint l;             /* nonnegative-cone rows, next */
  idxint *q;            /* second-order cone sizes, last, in this order */
  idxint qsize;
} Cone;

Sol *pdos(const Data *d, const Cone *k);
void freeData(Data **d, Cone **k);
void freeSol(Sol **sol);

/* CLOCK_MONOTONIC stopwatch used by the demo CLI (util.h:37-47). */
typedef struct timer {
  struct timespec tic;
  struct timespec toc;
} timer;
void tic(timer *t);
double toc(timer *t);   /* prints "time: %8.4f seconds." and returns seconds */
double tocq(timer *t);  /* returns seconds */

#ifdef __cplusplus
}
#endif

#endif /* PDOS_AMD_PDOS_H */
