// device.hpp -- records shared by host and device code.
#pragma once

#include <cstdint>

namespace pdos_amd {

// Result of one solve, written by the device.
struct DevResult {
  int iters;      // ADMM updates performed
  int status;     // 0 Solved, 1 Indeterminate
  int cg_iters;   // total CG iterations
  int fallback;   // 1 if a warm start was rejected for the cold start
  double p_res, d_res, p_obj, d_obj, eta;
};

// Byte offsets, in LDS, of the static data the resident variant of the
// workgroup engine (k_wg_res) keeps there for the whole solve: a group's
// image (copied in one coalesced pass) followed by the instance's b and c,
// its iterates and the reduction slots.
// Rows are processed in wave batches {first, count, log2 G, 0}: count rows of
// one class, G lanes each (G = the least power of two, at most 64, with at
// most 4 entries per lane), so no lane walks a long row alone.
// A product and A' product keep their rows in batch order: position q is
// *pos[q] = {row, first entry, end entry, 0} of *idx / *val (direct plugin:
// A's column indices also as pivot positions, ridp); the sweeps likewise.
struct ResLayout {
  int D, E, rpos, ridx, rval, ridp, rbat, cpos, cidx, cval, cbat, soc, perm, pinv;
  int fpos, fidx, fval, fbat, flb, bpos, bidx, bval, bdq, bbat, blb;
  int nrbat, ncbat;  // wave batches of A's rows and of A''s
  int mbat;          // most wave batches any one phase (an SpMV or a sweep level) runs
  int nbig, big;     // second-order cones of more than kWgBigCone rows: count and {r0, r1} pairs
  int nfl, nbl;      // levels of the two sweeps (T excluded, forward: plus T's outside entries)
  int t, trow, tm;   // the dense trailing chain: t rows, their ids, the inverse of its t x t L D L' block
                     // (row-major)
  int img;          // image bytes (multiple of 16)
  int b, c, vec;    // per-instance b, c (binary64) and the iterates
  int total;        // dynamic LDS bytes
};

// One problem instance for the workgroup engine.  All pointers are device
// pointers; matrix values have the engine's bulk type (double or float).
struct WgInst {
  int n, m, f, l, nsoc, N;
  int nrb, ncb;            // row blocks of A rows / A' rows (spmv_blocks)
  int max_iters, cg_max_its;
  int nfl, nbl;            // direct: forward / backward level counts
  double eps, alpha, cg_tol_sq, lambda, pscale, dscale;
  const int* soc_ptr;      // nsoc + 1 absolute row offsets
  const int* rp;           // A rows (CSR), m + 1
  const int* ri;
  const void* rx;
  const int* cp;           // A' rows = A columns (CSC), n + 1
  const int* ci;
  const void* cx;
  const int* rblk;         // A row blocks {r0, r1, p0, p1}, nrb of them (int4)
  const int* cblk;         // A' row blocks, ncb of them
  const double* b;         // scaled b, c and the scalings
  const double* c;
  const double* D;
  const double* E;
  const double* x0;        // scaled warm start (read once)
  const double* s0;
  const double* y0;
  const float* wsrc;        // optional warm start in original units: [x (n), y (m), s (m)]
  int warm_check;          // 1: compare the warm start's residuals with the cold start's
  int pad3;
  double* x;               // final scaled iterate (written once)
  double* s;
  double* y;
  // direct plugin
  const int* perm;         // N
  const int* lrp;          // L rows (forward)
  const int* lri;
  const double* lrx;
  const int* lcp;          // L columns (backward)
  const int* lci;
  const double* lcx;
  const double* ld;        // D of L D L'
  const int* fl_ptr;
  const int* fl_rows;
  const int* bl_ptr;
  const int* bl_rows;
  DevResult* res;
  int lds_bytes;           // dynamic LDS this instance needs
  int hist_rows;           // rows of hist (0: no history)
  double* hist;            // VERBOSE: residuals (p_res, d_res, p_obj, d_obj, eta) of checks 0, 10, 20, ...
  double* cg_hist;         // VERBOSE, indirect: per update {CG steps, final |q|} (hist_rows * 10 rows)
  // resident variant: matrices, factor and level sets in LDS
  int resident;            // kernel variant: 0 k_wg_solve, 1 k_wg_res (image in LDS), 2 k_wg_res (image in memory)
  int pad4;
  const unsigned char* img;  // the group's LDS image in device memory (lay.img bytes)
  ResLayout lay;
};

// Setup on the device (batch_setup.hip): one record per distinct matrix of a
// batch and one per instance.  All pointers are device pointers into the
// batch's buffer.
struct SetupGroup {
  int n, m, nnz, normalize, nsoc, pad;
  const int* cp;           // CSC column pointers
  const int* ci;           // CSC row indices (no image), or null
  const int* cidx;         // the image's row indices of A's columns (image), or null
  double* ax;              // the caller's values in CSC order, scaled in place (with the x1e3)
  const int* rp;           // row pattern of A
  const int* csr_src;      // CSC entry of each row entry
  const int* soc;          // nsoc + 1 absolute row offsets
  double* D;               // the image's D (m) and E (n, with the x1e3)
  double* E;
  double* E0;              // E before the x1e3 (n)
  double* D2;              // the streaming variant's D and E, or null
  double* E2;
  void* rval;              // image values of A's rows and columns (bulk type), or null (no image)
  void* cval;
  const int4* rpos;        // the image's positions {row, e0, e1, 0} of A's m rows and n columns
  const int4* cpos;
  void* srx;               // the streaming variant's values (bulk type) in row and column order, or null
  void* scx;
};

struct SetupInst {
  int n, m, rec, cls;      // record in instance order and in class order
  double* b;               // the caller's b, c and warm start, scaled in place
  double* c;
  double* x;
  double* s;
  double* y;
  const double* D;         // the group's D and E0
  const double* E0;
  double* out;             // lambda, pscale, dscale
};

// Rows of the VERBOSE residual history for a MAX_ITERS: one per tenth check
// (pdos.c:107 prints checks i % 10 == 0), at most 100,000.
inline int verbose_hist_rows(long max_iters) {
  const long r = (max_iters + 9) / 10;
  return (int)(r < 1 ? 1 : (r > 100000 ? 100000 : r));
}

constexpr int kWgThreads = 256;
constexpr int kWgBigCone = 256;       // cones with more rows are projected by the whole workgroup
constexpr int kWgChunk = 2048;        // entries per SpMV row block (products staged in LDS)
constexpr int kWgRowsPerBlock = 1024;  // rows per SpMV row block
constexpr int kWgLongRow = 64;        // rows with this many entries are summed by a wave

// Iterates of the resident variant, in LDS after b and c: x (n), s, y, s~,
// A x (m) in the bulk type; indirect: p, q, u (n), t (m); direct: the KKT
// vector (n + m binary64) and 64 binary64 of staging for the trailing
// block; then two reduction slots.
inline long res_vec_bytes(int n, int m, int N, bool direct, int tsize) {
  auto al = [](long v) { return (long)((v + 15) & ~15L); };
  long bytes = al((long)(n + 4 * (long)m) * tsize);
  if (direct) bytes += al((long)N * 8) + 64 * 8;  // the KKT vector; the trailing block's staging
  else bytes += al((long)(3 * (long)n + m) * tsize);
  return bytes + 2 * 16 * 4 * 8;
}

// Dynamic LDS bytes for one instance in the workgroup engine.
inline int wg_lds_bytes(int n, int m, int N, bool direct, int tsize) {
  auto al = [](long v) { return (long)((v + 15) & ~15L); };
  long bytes;
  if (!direct) bytes = al((long)(4 * (long)n + 5 * (long)m) * tsize);
  else bytes = al((long)(n + 4 * (long)m) * tsize) + al((long)N * 8);
  bytes += al((long)kWgChunk * tsize);  // SpMV products
  bytes += 2 * 16 * 4 * 8;  // reduction slots
  return (int)bytes;
}

}  // namespace pdos_amd
