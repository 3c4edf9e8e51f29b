/*
 * pdos_ext.h -- extensions of the MI355X PDOS library (no reference
 * counterpart; the reference API is pdos.h:85-87 only).  They never change the
 * layout of the reference structs.
 *
 *   pdos_ex            pdos() with explicit plugin / precision / engine and
 *                      a PdosInfo side channel (the reference's Sol carries no
 *                      iteration count or residuals, pdos.h:37-41)
 *   pdos_batch         many independent cone programs of uneven size in one
 *                      persistent launch (one workgroup per instance)
 *   pdos_family_*      a family of instances that share A and b (only c, or
 *                      b and c, vary): factor once, solve batches on MFMA
 *   pdos_last_error    text of the last failure on this thread
 *
 * Precision: PDOS_FP64 keeps every vector, SpMV and reduction in IEEE binary64
 * (parity mode).  PDOS_MIXED stores matrices and iterates in binary32 and
 * keeps the stopping norms, objectives and CG scalars in binary64.
 * pdos() reads PDOS_PRECISION=fp64|mixed from the environment (default fp64).
 */
#ifndef PDOS_AMD_PDOS_EXT_H
#define PDOS_AMD_PDOS_EXT_H

#include <stddef.h>

#include "pdos.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PDOS_INDIRECT = 0, PDOS_DIRECT = 1 };
enum { PDOS_FP64 = 0, PDOS_MIXED = 1 };
enum { PDOS_ENGINE_AUTO = 0, PDOS_ENGINE_WG = 1, PDOS_ENGINE_GRID = 2 };
/* Scaling: the reference build's (common.h:35-218), or the alternative header
 * the reference ships to be compiled in its place (ruiz_equil_common.h:98-322:
 * with NORMALIZE, three Ruiz-type sweeps over [A b; c' 0] with cone-collapsed
 * maxima; lambda = (1e-6 + |b|) / (1e-6 + |c|) in both cases).  pdos() takes
 * it from the environment: PDOS_EQUIL=ruiz. */
enum { PDOS_EQUIL_REFERENCE = 0, PDOS_EQUIL_RUIZ = 1 };

typedef struct PdosOptions {
  int plugin;      /* PDOS_INDIRECT / PDOS_DIRECT */
  int precision;   /* PDOS_FP64 / PDOS_MIXED */
  int engine;      /* PDOS_ENGINE_*: WG = one workgroup per instance,
                      GRID = whole-GPU kernels for one large instance */
  int device;      /* HIP device ordinal */
  int block;       /* reserved, 0 (the WG engine runs 256 threads per workgroup) */
  int bf16_products; /* family engine, mixed: products of the split-bf16 x-update
                        GEMM, 1 (hi*hi), 2 (+hi*lo) or 3 (+lo*hi); 0 = 1.  The
                        GEMM computes the correction x+ - x, whose error shrinks
                        with the step, so one product converges like three. */
  int equil;       /* PDOS_EQUIL_* (the family engine: reference only) */
  int tune;        /* batches (pdos_batch, pdos_batch_prepare): 1 = time candidate
                      launch shapes of the workgroup engine per size bucket of
                      n + m (kernel variant, persistent work-queue launch,
                      entries per lane, threads per workgroup) on samples
                      of the batch and use the
                      fastest; decisions are cached per process.  0 = the
                      built-in rules. */
} PdosOptions;

typedef struct PdosInfo {
  long iters;        /* ADMM updates performed (the reference's loop index) */
  long cg_iters;     /* total CG iterations (indirect) */
  long status;       /* 0 Solved, 1 Indeterminate */
  long engine;       /* engine that ran */
  long fallback;     /* 1 if the warm start was rejected for the cold start */
  double p_res, d_res, p_obj, d_obj, eta;  /* last stop check */
  double lambda;
  double setup_ms;   /* host setup + factorization + upload */
  double solve_ms;   /* device solve (host wall, synchronous) */
  double total_ms;
  double lnz;        /* nnz(L) for the direct plugin */
} PdosInfo;

/* Default options: the library's plugin (indirect in libpdosindir_hip.so,
 * direct in libpdosdir_hip.so), fp64, auto engine, device 0, the reference
 * scaling. */
void pdos_default_options(PdosOptions *opt);

/* pdos() with options and info (info may be NULL). */
Sol *pdos_ex(const Data *d, const Cone *k, const PdosOptions *opt, PdosInfo *info);

/* Info of the last pdos()/pdos_ex() on this thread; returns 0 if any. */
int pdos_last_info(PdosInfo *info);

/* Solves count independent problems; sols[i] receives a malloc'd Sol (free
 * each with freeSol), infos (optional) per-instance info.  Returns 0 or -1
 * (see pdos_last_error).  Setup and the direct plugin's factorizations run on
 * host threads (PDOS_HOST_THREADS, default the hardware threads, at most 64). */
int pdos_batch(const Data *const *ds, const Cone *const *ks, int count, const PdosOptions *opt,
               Sol **sols, PdosInfo *infos);

/* The same batch split in three so that callers (and benchmarks) can time the
 * device solve with every input already resident in device memory:
 *   prepare  host setup (+ factorization) and one upload
 *   run      the solve on the device from the prepared start; re-runnable,
 *            each run starts again from the inputs' warm start (or zero);
 *            *kernel_ms (optional) receives the HIP-event time of the launch
 *   fetch    download and unscale into malloc'd Sols (free with freeSol)
 * A handle belongs to the thread that prepared it. */
typedef struct PdosBatchHandle PdosBatchHandle;
PdosBatchHandle *pdos_batch_prepare(const Data *const *ds, const Cone *const *ks, int count,
                                    const PdosOptions *opt);
int pdos_batch_run(PdosBatchHandle *h, double *kernel_ms);
int pdos_batch_fetch(PdosBatchHandle *h, Sol **sols, PdosInfo *infos);
void pdos_batch_free(PdosBatchHandle *h);
/* Device memory a prepared batch holds (bytes; 0 for NULL). */
size_t pdos_batch_device_bytes(const PdosBatchHandle *h);
/* The tuner's decisions for a batch prepared with opt->tune (JSON: per size
 * bucket the n + m range, instances, sample size, chosen variant, launch mode,
 * entries per lane and threads per workgroup, and every candidate's
 * [variant, persistent, entries_per_lane, threads, kernel ms]); "" without
 * tuning.  Owned by the handle. */
const char *pdos_batch_tuning(const PdosBatchHandle *h);
/* Host wall time (ms) of pdos_batch_prepare's parts: ms[0] setup (host
 * threads: checks, the row pattern; scaling and transpose where the setup
 * stays on the host), ms[1] grouping by matrix and the direct plugin's
 * factorizations, ms[2] launch-shape tuning (0 without opt->tune), ms[3]
 * packing and upload.  Returns 0, or -1 for NULL arguments. */
int pdos_batch_prepare_times(const PdosBatchHandle *h, double *ms);
/* Scalings of instance i of a prepared batch (common.h:69-192): D (m), E (n,
 * with the x1e3) and lambda, wherever they were formed (the indirect
 * plugin's workgroup-engine members are set up on the device; the others on
 * the host).  Any output may be NULL.  Returns 1 if instance i was set up on
 * the device, 0 if on the host, -1 for bad arguments. */
int pdos_batch_scales(const PdosBatchHandle *h, int i, double *D, double *E, double *lambda);

/* Warm start for every instance of a prepared batch from device memory:
 * instance i reads [x (n), y (m), s (m)] in original units (float) at
 * dW + i * ld.  With check != 0 the solver compares, at its first stop test,
 * the warm start's scaled residuals max(p_res/(1+|b|), d_res/(1+|c|), gap)
 * with the cold start's and falls back to the cold start if they are worse
 * (PdosInfo.fallback).  dW == NULL restores the prepared start. */
int pdos_batch_set_warm_dev(PdosBatchHandle *h, const float *dW, long ld, int check);

/* Device memory for callers that keep inputs resident (plain HIP memory on
 * the library's device; no torch). */
void *pdos_device_alloc(size_t bytes);
void pdos_device_free(void *p);
int pdos_device_upload(void *dst, const void *src, size_t bytes);
int pdos_device_download(void *dst, const void *src, size_t bytes);
/* dst = src for device memory, by one plain copy kernel of this library (the
 * reference kernel for scaling memory-traffic counters); synchronous. */
int pdos_device_copy(void *dst, const void *src, size_t bytes);

/* A family of cone programs sharing A, the cones and the parameters of the
 * template data d (its b is the default b, its c is ignored), differing in c
 * (and optionally b).  Solved in lockstep with G = (I + A'A)^-1 formed once:
 * the x-update is one GEMM for the whole batch (PDOS_FP64: rocBLAS dgemm;
 * PDOS_MIXED: the correction x + G (A'W - lambda c) with one bf16 MFMA
 * product accumulated in binary32; rows of A with 64 or more entries are
 * multiplied as binary32 MFMA tiles).  A mixed load of 8,192 or more
 * instances is split over two engines that run at once.
 *   load   C: count x n, B: count x m or NULL (original units, row-major)
 *   run    warm_dev: optional device array count x ld of [x, y, s] in original
 *          units; judge != 0 compares it with the cold start per instance
 *   fetch  X: count x n, Y, S: count x m (any may be NULL), iters, status
 *          (0 Solved, 1 Indeterminate), fallback, resid (count x 5: p_res,
 *          d_res, c'x, -b'y, gap of the last stop check) */
typedef struct PdosFamily PdosFamily;
PdosFamily *pdos_family_create(const Data *d, const Cone *k, const PdosOptions *opt);
int pdos_family_load(PdosFamily *h, int count, const double *C, const double *B);
int pdos_family_run(PdosFamily *h, const float *warm_dev, long ld, int judge, double *kernel_ms);
int pdos_family_fetch(PdosFamily *h, double *X, double *Y, double *S, long *iters, long *status,
                      long *fallback, double *resid);
void pdos_family_free(PdosFamily *h);
const char *pdos_family_last_error(void);
/* Per-kernel HIP-event timing of the solve loop.  Returns (and resets) the
 * totals since the last call in ms[6] / calls[6] / cols[6] (instance columns
 * the launches covered, padding included) for the kinds: phase A, A' product,
 * phase B, x-update GEMM, A product, phase C; then turns the timing on or off
 * for later runs (events add a little launch overhead). */
int pdos_family_timing(PdosFamily *h, int enable, double *ms, long *calls, double *cols);
/* Engines the next pdos_family_load splits a large mixed-precision batch
 * over (1 or 2; n <= 0 only queries); returns the setting. */
int pdos_family_engines(PdosFamily *h, int n);
/* Per loaded instance (the order of pdos_family_load): lambda, and the stop
 * test's scales 1 + |b| and 1 + |c| (pdos.c:76-77), as the device computed
 * them at load (common.h:182-205 on the instance's b and c); any may be NULL. */
int pdos_family_scales(PdosFamily *h, double *lam, double *pscale, double *dscale);
/* out[9]: n, m, padded n of the GEMM, compactions in the last run, dense A,
 * fp64, bf16 products of the mixed x-update, dense rows of A multiplied as
 * MFMA tiles (0: none), engines the loaded batch is split over (1 or 2). */
int pdos_family_info(PdosFamily *h, long *out);

/* Learned warm start: MLP n_in -> n_hidden -> n_hidden -> n_out with GELU,
 * bf16 weights mirrored from a binary32 master, binary32 accumulation, all
 * products on BF16 MFMA (hand-written kernels), Adam training.
 * Parameters (get/set, gradients) use the layout W1 [n_in x n_hidden],
 * b1, W2 [n_hidden x n_hidden], b2, W3 [n_hidden x n_out], b3, row-major.
 * Inputs pass through (x - in_mean) / in_std; outputs of predict are
 * y * out_std + out_mean; training targets are in the standardised space. */
typedef struct PdosMlp PdosMlp;
PdosMlp *pdos_mlp_create(int n_in, int n_hidden, int n_out, unsigned long seed, int device);
void pdos_mlp_free(PdosMlp *P);
int pdos_mlp_num_params(const PdosMlp *P);
int pdos_mlp_get_params(PdosMlp *P, float *out);
int pdos_mlp_set_params(PdosMlp *P, const float *in);
int pdos_mlp_set_affine(PdosMlp *P, const float *in_mean, const float *in_std, const float *out_std,
                        const float *out_mean);
int pdos_mlp_loss_grad(PdosMlp *P, const float *X, const float *T, int rows, double *loss,
                       float *grads);
int pdos_mlp_train(PdosMlp *P, const float *X, const float *T, int rows, int epochs, int batch,
                   float lr, unsigned long seed, double *last_loss);
int pdos_mlp_predict(PdosMlp *P, const float *X, int rows, float *Y, double *kernel_ms);
int pdos_mlp_predict_dev(PdosMlp *P, const float *dX, int rows, float *dY, int ldy);
/* pdos_mlp_predict_dev that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * The forward pass's activations live in a fixed 32 MiB device array of the
 * library, one per GPU (held until the process ends); rows are processed in
 * chunks that fit it (2,176 rows for the 501-512-512-2545 network).  A
 * network whose 64 rows do not fit is refused (-1, pdos_mlp_last_error).
 * Results equal pdos_mlp_predict_dev's bit for bit.  Calls that overlap on
 * one GPU (several threads or networks) race on the array. */
int pdos_mlp_predict_dev_noalloc(PdosMlp *P, const float *dX, int rows, float *dY, int ldy);
const char *pdos_mlp_last_error(void);

/* Host-side check of the direct plugin's factorization: builds the scaled
 * KKT [[-I, A'], [A, I]] of (d, k) exactly as pdos() does, orders and factors
 * it, and solves K z = rhs (length n+m) on the host with the same level-set
 * schedule the GPU sweeps use.  stats (optional, 4 longs): N, nnz(L),
 * forward levels, backward levels.  Returns 0 or -1. */
int pdos_kkt_solve_host(const Data *d, const Cone *k, const double *rhs, double *z, long *stats);

/* The direct plugin's fill-reducing ordering of the scaled KKT of (d, k):
 * perm (n+m longs) receives the pivot order (perm[k] = original index of the
 * k-th pivot).  Returns 0 or -1. */
int pdos_kkt_order(const Data *d, const Cone *k, long *perm);

/* Host-side check of the setup: the scaling the solver iterates with, as the
 * reference's initWork leaves it in Work (common.h:35-218, or with
 * equil = PDOS_EQUIL_RUIZ ruiz_equil_common.h:98-322): D (m), E (n, with the
 * x1e3) and lambda.  Returns 0 or -1 (pdos_last_error). */
int pdos_setup_host(const Data *d, const Cone *k, int equil, double *D, double *E, double *lambda);

/* Text of the last error on this thread ("" if none). */
const char *pdos_last_error(void);

/* Library build string. */
const char *pdos_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PDOS_AMD_PDOS_EXT_H */
