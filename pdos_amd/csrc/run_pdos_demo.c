/*
 * run_pdos_demo.c -- the demo CLI of the reference (run_pdos.c:9-115) on the
 * MI355X library: reads one problem file in the demo's binary format, calls
 * pdos(), prints the timings, frees everything with freeData / freeSol.
 *
 *   demo_indirect_hip [problem-file]    (linked against libpdosindir_hip.so)
 *   demo_direct_hip   [problem-file]    (linked against libpdosdir_hip.so)
 *
 * File format (run_pdos.c:31-94), little-endian, idxint = int64, double:
 *   n m f l qsize | q[qsize] | b[m] | c[n] | MAX_ITERS CG_MAX_ITS |
 *   ALPHA EPS_ABS CG_TOL | VERBOSE NORMALIZE | Anz | Ai[Anz] Ap[n+1] Ax[Anz]
 * Warm starts are not part of the format (Data.x = y = s = NULL).
 */
#include <stdio.h>
#include <stdlib.h>

#include "../../include/pdos.h"

static int read_all(FILE *fp, void *dst, size_t size, size_t count) {
  return fread(dst, size, count, fp) == count ? 0 : -1;
}

static int read_problem(FILE *fp, Data **dp, Cone **kp) {
  Data *d = (Data *)calloc(1, sizeof(Data));
  Cone *k = (Cone *)calloc(1, sizeof(Cone));
  idxint anz = 0;
  int bad = 0;
  d->p = (Params *)calloc(1, sizeof(Params));
  bad |= read_all(fp, &d->n, sizeof(idxint), 1);
  bad |= read_all(fp, &d->m, sizeof(idxint), 1);
  bad |= read_all(fp, &k->f, sizeof(idxint), 1);
  bad |= read_all(fp, &k->l, sizeof(idxint), 1);
  bad |= read_all(fp, &k->qsize, sizeof(idxint), 1);
  if (bad || d->n < 1 || d->m < 1 || k->qsize < 0) goto fail;
  k->q = (idxint *)malloc(sizeof(idxint) * (size_t)(k->qsize + 1));
  bad |= read_all(fp, k->q, sizeof(idxint), (size_t)k->qsize);
  d->b = (double *)malloc(sizeof(double) * (size_t)d->m);
  d->c = (double *)malloc(sizeof(double) * (size_t)d->n);
  bad |= read_all(fp, d->b, sizeof(double), (size_t)d->m);
  bad |= read_all(fp, d->c, sizeof(double), (size_t)d->n);
  bad |= read_all(fp, &d->p->MAX_ITERS, sizeof(idxint), 1);
  bad |= read_all(fp, &d->p->CG_MAX_ITS, sizeof(idxint), 1);
  bad |= read_all(fp, &d->p->ALPHA, sizeof(double), 1);
  bad |= read_all(fp, &d->p->EPS_ABS, sizeof(double), 1);
  bad |= read_all(fp, &d->p->CG_TOL, sizeof(double), 1);
  bad |= read_all(fp, &d->p->VERBOSE, sizeof(idxint), 1);
  bad |= read_all(fp, &d->p->NORMALIZE, sizeof(idxint), 1);
  bad |= read_all(fp, &anz, sizeof(idxint), 1);
  if (bad || anz < 0) goto fail;
  d->Ai = (idxint *)malloc(sizeof(idxint) * (size_t)(anz + 1));
  d->Ap = (idxint *)malloc(sizeof(idxint) * (size_t)(d->n + 1));
  d->Ax = (double *)malloc(sizeof(double) * (size_t)(anz + 1));
  bad |= read_all(fp, d->Ai, sizeof(idxint), (size_t)anz);
  bad |= read_all(fp, d->Ap, sizeof(idxint), (size_t)(d->n + 1));
  bad |= read_all(fp, d->Ax, sizeof(double), (size_t)anz);
  if (bad) goto fail;
  *dp = d;
  *kp = k;
  return 0;
fail:
  freeData(&d, &k);
  return -1;
}

int main(int argc, char **argv) {
  timer t;
  Data *d = NULL;
  Cone *k = NULL;
  Sol *sol;
  FILE *fp;
  int rc;
  if (argc < 2) {
    fprintf(stderr, "usage: %s problem-file\n", argv[0]);
    return 2;
  }
  tic(&t);
  fp = fopen(argv[1], "rb");
  if (!fp) {
    fprintf(stderr, "couldn't open %s\n", argv[1]);
    return 1;
  }
  if (read_problem(fp, &d, &k) != 0) {
    fclose(fp);
    fprintf(stderr, "%s: not a problem file\n", argv[1]);
    return 1;
  }
  fclose(fp);
  printf("File IO %4f seconds\n", tocq(&t));
  tic(&t);
  sol = pdos(d, k);
  printf("Total factorize + solve time %4f seconds\n", tocq(&t));
  rc = sol ? 0 : 1;
  if (sol) printf("status: %s\n", sol->status);
  freeData(&d, &k);
  freeSol(&sol);
  return rc;
}
