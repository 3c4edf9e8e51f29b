// family_engine.hip -- many cone programs that share A (and the cones), each
// with its own c (and optionally b), solved in lockstep.
//
// Every per-instance vector is stored feature-major, X[row * ldb + instance],
// so a wave of 64 lanes holds 64 instances: the shared matrix entries are
// wave-uniform (scalar loads, broadcast) and each instance's values are one
// coalesced 256-byte (fp32) row segment.  With G = (I + A'A)^-1 formed once,
// the x-update of the direct plugin (direct/private.c:16-56) is one GEMM for
// the whole batch:
//
//   SpMM      U = A'Y, V = A'W                                   (one pass over A')
//   phase B   |(c + U) ./ E|^2, c'X, R = X - lam c + V (fp64) or V - lam c (mixed)
//   decide    stop test per instance
//   GEMM      XN = G R     (fp64: rocBLAS dgemm)
//             XN = X + G R (mixed: bf16 MFMA; the correction form keeps the bf16
//             error proportional to the step)
//   SpMM      AX = A XN
//   phase C   X = XN; S~ = b - AX; relax; S = Pi_K(S~ - lam Y); Y += (S - S~)/lam;
//             then, for the next iteration's A' product and stop test, W and
//             the partials |(S - b + AX) ./ D|^2, b'Y per instance, where
//             W = b - S - lam Y (fp64) or b - AX - S - lam Y (mixed)
//
// Phase A forms W and those partials from scratch; it runs only when the
// iterate or the layout changed outside the loop (start, warm-start judge,
// and compaction unless A has dense rows).
//
// Mixed precision with dense rows in A (rows of 64 or more entries, at most
// 32): an iteration is four launches.  The A' product multiplies the dense
// rows as an fp32 MFMA tile and gathers the rest (k_fam_at_hyb); the
// x-update also forms the dense rows' part of A XN as one plane per row tile
// and runs the stop test in extra workgroups (k_fam_xupdate); phase C sums
// the planes and gathers the other rows' entries from XN itself, without
// storing AX, in two passes over the cone rows (k_fam_c1 / k_fam_c2, or the
// latency-first k_fam_c1_ll / k_fam_c2_ll for narrow batches).  Compaction
// then moves W and the stop-test partials instead of AX.
//
// An instance that meets the stop test (pdos.c:95-107) or MAX_ITERS is frozen:
// phases A-C skip it, so its iterate is exactly the one the stop test saw.
// Warm starts (original units) are scaled in on entry; with the judge on, an
// extra A/B pass compares the warm start's residuals with the cold start's and
// resets rejected instances to zero.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <type_traits>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/pdos_ext.h"
#include "engine.hpp"

namespace pdos_amd {

namespace {

constexpr int LANES = 64;  // instances per block
constexpr int NWV = 8;     // waves per block, splitting rows
constexpr int FT = LANES * NWV;

template <typename T>
struct FamArgs {
  int n, m, f, lend, nsoc, ldb, count, np, b_inst;
  int max_iters, judge;
  int nprod;           // mixed: planes of Rcat the x-update GEMM reads (1-3)
  double eps, relax;
  const int* soc_ptr;
  const int* rp; const int* ri; const T* rx;   // A rows
  const int* cp; const int* ci; const T* cx;   // A' rows
  const double* D; const double* E;
  T* b;                // m (shared) or m x ldb
  T* c;                // n x ldb, scaled (x1e3)
  double* lam; double* pscale; double* dscale;
  int* orig;           // column -> instance of the loaded batch
  double* pa;          // phase A partials [chunk][ldb][2]
  double* pb;          // phase B partials [chunk][ldb][2]
  double* pc;          // phase C cone-chunk partial norms [item][ldb]
  double* pt0;         // phase C first entry of each cone [cone][ldb]
  int nx_items;        // phase C: items of x rows (come first); pa index = item - nx_items
  int nlin_items;      // phase C: items of zero / nonneg rows (then the cone chunks)
  T* X; T* XN; T* S; T* Y; T* AX; T* W; T* U; T* V; T* R;
  __bf16* Rcat;        // mixed: [D_hi; D_lo; D_hi] of the correction D, 3 np x ldb
  // mixed, A with dense rows: phase C forms AX itself: per (item, wave, k)
  // the row's record {column of its only entry, bits of its value} or {-1,
  // the row's slot among the dense rows} (the sum of the x-update's planes
  // AXp, mt of them) / {-2 no entries} / {-3 several entries: the light CSR
  // lrp, lri, lrx}
  const int2* rec8; const float* AXp; int mt;
  const int* lrp; const int* lri; const float* lrx;
  int* done; int* status; int* iters; int* fallback;
  double* red;         // per instance: pres2, by, pres, dres, pobj, dobj, eta, spare
  // dense A, mixed: U and V arrive as nsplit planes of partial sums over K
  // (plane stride in elements), added by phase B in plane order
  int nsplit;
  long plane;
};

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Element of a feature-major row at a lane's byte offset: the row start is
// wave-uniform (computed on the scalar unit), the 32-bit offset per lane, so
// the access needs no 64-bit vector address arithmetic.
template <typename P>
__device__ __forceinline__ P& at_lane(P* row, unsigned boff) {
  return *reinterpret_cast<P*>(reinterpret_cast<char*>(row) + boff);
}
template <typename P>
__device__ __forceinline__ const P& at_lane(const P* row, unsigned boff) {
  return *reinterpret_cast<const P*>(reinterpret_cast<const char*>(row) + boff);
}

// The x-update right-hand side Rcat (3 np rows x ldb instances, bf16) is
// stored in blocks of 8 rows: element (k, inst) at (k / 8) * 8 ldb + 8 inst +
// k % 8, so the 8 values of one block for one instance -- an MFMA operand of
// 16x16x32 -- are 16 contiguous bytes.
__device__ __forceinline__ __bf16& rcat_at(__bf16* R, int ldb, int k, int inst) {
  return at_lane(R + ((size_t)(k >> 3) * ldb) * 8 + (k & 7), (unsigned)inst * 16u);
}

// Wave-uniform read of data no kernel writes while it runs (per-row scales,
// index tables): through the constant address space, so it is a scalar load
// waited on by itself rather than behind the vector loads in flight.
template <typename P>
__device__ __forceinline__ P ld_uniform(const P* p) {
  typedef const __attribute__((address_space(4))) P* cptr;
  return *(cptr)p;
}
__device__ __forceinline__ int2 ld_uniform(const int2* p) {
  const int* q = reinterpret_cast<const int*>(p);
  return int2{ld_uniform(q), ld_uniform(q + 1)};
}
__device__ __forceinline__ int4 ld_uniform(const int4* p) {
  const int* q = reinterpret_cast<const int*>(p);
  return int4{ld_uniform(q), ld_uniform(q + 1), ld_uniform(q + 2), ld_uniform(q + 3)};
}

// Sums one value per (wave, lane) over the waves of the block: every thread
// gets its lane's total.  Two barriers; `buf` is NWV x LANES doubles.
__device__ __forceinline__ double waves_sum(double v, double* buf) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  buf[w * LANES + lane] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NWV; ++k) s += buf[k * LANES + lane];
  __syncthreads();
  return s;
}

// As waves_sum for two values at once (one pair of barriers, each value
// summed in the same order); `buf` is 2 x NWV x LANES doubles.
__device__ __forceinline__ void waves_sum2(double& a, double& b, double* buf) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double2* b2 = reinterpret_cast<double2*>(buf);
  b2[w * LANES + lane] = double2{a, b};
  __syncthreads();
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int k = 0; k < NWV; ++k) {
    const double2 v = b2[k * LANES + lane];
    sa += v.x;
    sb += v.y;
  }
  __syncthreads();
  a = sa;
  b = sb;
}

// Mixed precision (binary32 bulk) uses the correction form of the x-update.
template <typename T>
constexpr bool kCorr = std::is_same<T, float>::value;

// The next A' right-hand side of row j from the finished (s, y):
//   fp64:  W = b - (s + lam y)
//   mixed: W = (b - A x) - (s + lam y), so A'W = A'(b - s - lam y) - A'A x.
template <typename T>
__device__ __forceinline__ T next_w(T b, T ax, T s, T y, T lam) {
  return kCorr<T> ? (b - ax) - (s + lam * y) : b - (s + lam * y);
}

template <typename T>
__device__ __forceinline__ T bval(const FamArgs<T>& a, int j, int inst) {
  return a.b_inst ? a.b[(size_t)j * a.ldb + inst] : a.b[j];
}

// Initial iterate from a warm start in original units (x/E, y/D, s*D) or zero.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_init(FamArgs<T> a, const float* __restrict__ warm, long wld) {
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const bool live = inst < a.count;
  for (int i = w; i < a.n; i += NWV) {
    T v = T(0);
    if (live && warm) v = T(double(warm[(size_t)inst * wld + i]) / a.E[i]);
    a.X[(size_t)i * a.ldb + inst] = v;
  }
  for (int r = a.n + w; r < a.np; r += NWV) {  // padding rows of the GEMM operands
    a.X[(size_t)r * a.ldb + inst] = T(0);
    if (a.Rcat)
      for (int k = 0; k < 3; ++k) rcat_at(a.Rcat, a.ldb, k * a.np + r, inst) = (__bf16)0.f;
  }
  for (int j = w; j < a.m; j += NWV) {
    T yv = T(0), sv = T(0);
    if (live && warm) {
      yv = T(double(warm[(size_t)inst * wld + a.n + j]) / a.D[j]);
      sv = T(double(warm[(size_t)inst * wld + a.n + a.m + j]) * a.D[j]);
    }
    a.Y[(size_t)j * a.ldb + inst] = yv;
    a.S[(size_t)j * a.ldb + inst] = sv;
  }
  if (w == 0) {
    a.done[inst] = live ? 0 : 1;
    a.status[inst] = 1;
    a.iters[inst] = 0;
    a.fallback[inst] = 0;
    for (int k = 0; k < 8; ++k) a.red[(size_t)inst * 8 + k] = -1.0;
  }
}

// Wave-uniform read of lane l's value (l uniform).
__device__ __forceinline__ int lane_of(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float lane_of(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_of(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const long long hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double((hi << 32) | lo);
}

constexpr int kUnroll = 8;  // gathers in flight per wave (measured 4 / 8 / 12 / 16 / 32: 121.4k / 121.5k / 120.4k / 120.0k / 112.6k SOCPs/s on the portfolio family)
// Mixed precision, sparse A: rows of A with at least kDenseRow entries (the
// kDenseMax longest) are the "dense rows" whose A' part is an fp32 MFMA tile
constexpr int kDenseRow = 64, kDenseMax = 32;
struct DenseRows {
  int row[kDenseMax];  // rows of A (unused slots: row 0, with zero coefficients)
};

template <typename T, int NR>
struct SpmmIn {
  const int4* tasks; const int* rowof; const int* idx; const T* val;
  const T* X1; const T* X2; const T* X3;
  int ldb;
  int cols, nitems;  // grid: cols column blocks x nitems task blocks, 1-D
};

// (column block bx, task block by) of this workgroup in a 1-D grid of
// cols * nitems.  The hardware deals workgroups to the 8 XCDs round robin;
// with cols % 8 == 0 each XCD's consecutive workgroups take all task blocks
// of one column block before the next, so the rows they gather for those 64
// instances stay in that XCD's L2.
__device__ __forceinline__ void spmm_block(int cols, int nitems, int& bx, int& by, int L = blockIdx.x) {
  if ((cols & 7) == 0) {
    const int seq = L >> 3;
    bx = (seq / nitems) * 8 + (L & 7);
    by = seq % nitems;
  } else {
    bx = L % cols;
    by = L / cols;
  }
}

// The wave's task of task block by.  begin(row) is called when the wave
// starts a row (before its entries are gathered: loads issued there overlap
// the row's products), done(row, acc) when the row is complete; for a split
// row, wave 0 gets done(row, block total).  Rows without entries inside a
// task are begun and done with zero sums.
// A wave's task and the (column, value, row) of its first 64 entries, loaded
// ahead by a caller that has other work to overlap them with.
template <typename T>
struct SpmmPre {
  int4 tk;
  int ci, ri;
  T vi;
  bool split;
};

template <typename T, int NR>
__device__ __forceinline__ SpmmPre<T> spmm_preload(const SpmmIn<T, NR>& in, int by) {
  const int lane = threadIdx.x & 63, w = wave_id();
  SpmmPre<T> p;
  p.tk = in.tasks[by * NWV + w];
  p.split = in.tasks[by * NWV].y < 0;
  p.ci = p.ri = 0;
  p.vi = T(0);
  const int q = p.tk.z, cnt = min(64, p.tk.w - q);
  if (lane < cnt) {
    p.ci = in.idx[q + lane];
    p.vi = in.val[q + lane];
    p.ri = p.split ? p.tk.x : in.rowof[q + lane];
  }
  return p;
}

template <typename T, int NR, typename Begin, typename Done>
__device__ __forceinline__ void spmm_core(const SpmmIn<T, NR>& in, int bx, int by, double* buf, Begin begin,
                                          Done done, const SpmmPre<T>* pre = nullptr) {
  const int lane = threadIdx.x & 63, w = wave_id();
  const unsigned boff = (unsigned)(bx * LANES + lane) * (unsigned)sizeof(T);
  const int4 tk = pre ? pre->tk : in.tasks[by * NWV + w];
  const bool split = pre ? pre->split : in.tasks[by * NWV].y < 0;  // block-uniform
  T acc[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) acc[k] = T(0);
  int cur = tk.x;
  if (split ? w == 0 : tk.x < tk.y) begin(cur);
  for (int q = tk.z; q < tk.w; q += 64) {
    const int cnt = min(64, tk.w - q);
    int ci = 0, ri = 0;
    T vi = T(0);
    if (pre && q == tk.z) {
      ci = pre->ci;
      vi = pre->vi;
      ri = pre->ri;
    } else if (lane < cnt) {
      ci = in.idx[q + lane];
      vi = in.val[q + lane];
      ri = split ? tk.x : in.rowof[q + lane];
    }
    for (int e0 = 0; e0 < cnt; e0 += kUnroll) {
      // entries past cnt repeat entry cnt - 1: in bounds, not accumulated
      T xv[kUnroll][NR], vv[kUnroll];
      int cc[kUnroll], rr[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int e = min(e0 + u, cnt - 1);
        cc[u] = lane_of(ci, e);
        vv[u] = lane_of(vi, e);
        rr[u] = lane_of(ri, e);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const size_t ro = (size_t)cc[u] * in.ldb;
        xv[u][0] = at_lane(in.X1 + ro, boff);
        if constexpr (NR > 1) xv[u][1] = at_lane(in.X2 + ro, boff);
        if constexpr (NR > 2) xv[u][2] = at_lane(in.X3 + ro, boff);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        if (e0 + u < cnt) {
          while (rr[u] != cur) {  // wave-uniform: rows up to rr[u] are complete
            done(cur, acc);
#pragma unroll
            for (int k = 0; k < NR; ++k) acc[k] = T(0);
            ++cur;
            begin(cur);
          }
#pragma unroll
          for (int k = 0; k < NR; ++k) acc[k] += vv[u] * xv[u][k];
        }
      }
    }
  }
  if (!split) {
    if (tk.x < tk.y) {
      for (;;) {
        done(cur, acc);
        if (++cur >= tk.y) break;
#pragma unroll
        for (int k = 0; k < NR; ++k) acc[k] = T(0);
        begin(cur);
      }
    }
    return;
  }
  T tot[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) tot[k] = T(waves_sum(double(acc[k]), buf));
  if (w == 0) done(tk.x, tot);
}

// Y_k = M X_k (k < NR) for a shared sparse M in CSR (idx, val, row of each
// entry) and feature-major X_k, up to three right-hand sides per pass.
// Work is host-built per wave: task {r0, r1, q0, q1} computes the complete
// rows [r0, r1) from the entries [q0, q1) (at most 64 of them); a long row is
// split over the 8 waves of its block, every task {r, -1, q0, q1}, and the
// block sums the waves' partials.  A wave loads up to 64 entries' (column,
// value, row) in one coalesced load per array, then issues the gathers of
// kUnroll entries before using any of them, so a task costs a few memory
// latencies however short its rows are.  Each row is summed in entry order.
template <typename T, int NR>
__global__ __launch_bounds__(FT) void k_fam_spmm(SpmmIn<T, NR> in, T* __restrict__ Y1, T* __restrict__ Y2,
                                                 T* __restrict__ Y3) {
  __shared__ double buf[2 * NWV * LANES];
  int bx, by;
  spmm_block(in.cols, in.nitems, bx, by);
  const unsigned boff = (unsigned)(bx * LANES + (threadIdx.x & 63)) * (unsigned)sizeof(T);
  spmm_core<T, NR>(in, bx, by, buf, [](int) {}, [&](int row, const T (&v)[NR]) {
    const size_t ro = (size_t)row * in.ldb;
    at_lane(Y1 + ro, boff) = v[0];
    if constexpr (NR > 1) at_lane(Y2 + ro, boff) = v[1];
    if constexpr (NR > 2) at_lane(Y3 + ro, boff) = v[2];
  });
}

// The A' product with phase B as its epilogue (sparse A): per row i of A',
// U_i = (A'Y)_i and V_i = (A'W)_i stay in registers and give the partial
// |(c + U) ./ E|^2 and c'X and the x-update right-hand side (as k_fam_b).
// c_i and X_i are loaded when the row begins.  Partials per block into
// pb[task block].
// Phase B for row i of instance inst, given U_i = u and V_i = v: the
// partials |(c + U) ./ E|^2 and c'X, and the x-update right-hand side.
template <typename T>
__device__ __forceinline__ void at_b_row(const FamArgs<T>& a, int i, int inst, T lam, T c_i, T x_i, double e_i,
                                         T u, T v, double& d2, double& cx) {
  const T r = (c_i + u) / T(e_i);
  d2 += double(r) * double(r);
  cx += double(c_i) * double(x_i);
  if (a.judge) return;
  const T rv = kCorr<T> ? v - lam * c_i : x_i - lam * c_i + v;
  if (kCorr<T>) {
    const __bf16 hi = (__bf16)float(rv);
    rcat_at(a.Rcat, a.ldb, i, inst) = hi;
    if (a.nprod > 1) rcat_at(a.Rcat, a.ldb, a.np + i, inst) = (__bf16)(float(rv) - float(hi));
    if (a.nprod > 2) rcat_at(a.Rcat, a.ldb, 2 * a.np + i, inst) = hi;
  } else {
    at_lane(a.R + (size_t)i * a.ldb, (unsigned)inst * (unsigned)sizeof(T)) = rv;
  }
}

// at_b_row's arithmetic for the mixed path, returning the right-hand side
// value instead of storing it.
template <typename T>
__device__ __forceinline__ T at_b_rv(const FamArgs<T>& a, int i, T lam, T c_i, T x_i, double e_i, T u, T v,
                                     double& d2, double& cx) {
  const T r = (c_i + u) / T(e_i);
  d2 += double(r) * double(r);
  cx += double(c_i) * double(x_i);
  return v - lam * c_i;
}

template <typename T>
__device__ __forceinline__ void at_b_partials(const FamArgs<T>& a, int by, int inst, bool act, double d2,
                                              double cx, double* buf) {
  waves_sum2(d2, cx, buf);
  if (act && wave_id() == 0) {
    double* pb = a.pb + ((size_t)by * a.ldb + inst) * 2;
    pb[0] = d2;
    pb[1] = cx;
  }
}

template <typename T>
__global__ __launch_bounds__(FT) void k_fam_at_b(FamArgs<T> a, SpmmIn<T, 2> in) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63;
  int bx, by;
  spmm_block(in.cols, in.nitems, bx, by);
  const int inst = bx * LANES + lane;
  const bool act = inst < a.count && !a.done[inst];
  const T lam = T(a.lam[inst]);
  T c_i = T(0), x_i = T(0);
  double e_i = 1.0;  // converted where used: the scalar load is waited for there
  double d2 = 0.0, cx = 0.0;
  const unsigned boff = (unsigned)inst * (unsigned)sizeof(T);
  spmm_core<T, 2>(in, bx, by, buf,
                  [&](int i) {
                    const size_t ro = (size_t)i * a.ldb;
                    c_i = at_lane(a.c + ro, boff);
                    x_i = at_lane(a.X + ro, boff);
                    e_i = ld_uniform(a.E + i);
                  },
                  [&](int i, const T (&v)[2]) {
                    if (act) at_b_row(a, i, inst, lam, c_i, x_i, e_i, v[0], v[1], d2, cx);
                  });
  at_b_partials(a, by, inst, act, d2, cx, buf);
}

typedef float hf32x4 __attribute__((ext_vector_type(4)));

// The A' product with phase B (mixed precision) for an A with dense rows.
// Workgroup: a 64 x 64 tile (rows m0.. of A', instances n0..).  The dense
// rows' part U_h = A'_h Y_h, V_h = A'_h W_h (K = kDenseMax rows of Y and W)
// runs on the fp32 MFMA 16x16x4, each wave two 16 x 16 sub-tiles of each,
// and the tiles meet in LDS; then wave w takes rows m0 + 8 w .. + 7: their
// other entries (the light CSR `in`, one task per wave and tile) are gathered
// and added, and the row epilogue is k_fam_at_b's.  Both parts are summed in
// binary32 before the right-hand side is rounded to bf16.  Ah holds A'_h in
// MFMA operand order: 64 floats per (16-row group g, K quad s) at
// (g * 8 + s) * 64 + lane, lane holding A'_h[16 g + lane % 16][4 s + lane / 16].
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_at_hyb(FamArgs<T> a, SpmmIn<T, 2> in, const float* __restrict__ Ah,
                                                   DenseRows dr) {
  __shared__ double buf[2 * NWV * LANES];
  __shared__ float Ut[64][65], Vt[64][65];
  __shared__ unsigned short rsh[NWV][8][LANES];  // the wave's 8 rows of Rcat (bf16 bits), stored as one block
  const int lane = threadIdx.x & 63, w = wave_id();
  int bx, ti;
  spmm_block(in.cols, in.nitems, bx, ti);
  const int m0 = ti * 64, n0 = bx * LANES;
  const SpmmPre<T> pre = spmm_preload(in, ti);  // in flight with the dense tile's operands
  // the instance's flag and lambda, also in flight (no load waits behind the barrier)
  const int inst = n0 + lane;
  const int dn = a.done[inst];
  const double lam_d = a.lam[inst];
  {
    const int rg = w >> 1, cg0 = (w & 1) * 2, kq = lane >> 4, col = lane & 15;
    float av[8], yb[2][8], wb[2][8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      av[s] = Ah[((size_t)(ti * 4 + rg) * 8 + s) * 64 + lane];
      const int r0 = dr.row[4 * s], r1 = dr.row[4 * s + 1], r2 = dr.row[4 * s + 2], r3 = dr.row[4 * s + 3];
      const int r = kq == 0 ? r0 : (kq == 1 ? r1 : (kq == 2 ? r2 : r3));
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const size_t o = (size_t)r * a.ldb + n0 + (cg0 + c) * 16 + col;
        yb[c][s] = float(a.Y[o]);
        wb[c][s] = float(a.W[o]);
      }
    }
    hf32x4 au[2], aw[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      au[c] = hf32x4{0.f, 0.f, 0.f, 0.f};
      aw[c] = hf32x4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        au[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], yb[c][s], au[c], 0, 0, 0);
        aw[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[s], wb[c][s], aw[c], 0, 0, 0);
      }
    // accumulator lane l: rows (l / 16) * 4 + r, column l % 16
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Ut[rg * 16 + kq * 4 + r][(cg0 + c) * 16 + col] = au[c][r];
        Vt[rg * 16 + kq * 4 + r][(cg0 + c) * 16 + col] = aw[c][r];
      }
  }
  __syncthreads();
  const bool act = inst < a.count && !dn;
  const T lam = T(lam_d);
  T c_i = T(0), x_i = T(0), uh = T(0), vh = T(0);
  double e_i = 1.0, d2 = 0.0, cx = 0.0;
  const unsigned boff = (unsigned)inst * (unsigned)sizeof(T);
  // one bf16 product and not the judge's pass: the wave's task is rows
  // rbase .. rbase + 7 (host-built so), whose right-hand side is stored at once
  const bool one = !a.judge && a.nprod == 1;
  const int rbase = m0 + 8 * w;
  spmm_core<T, 2>(in, bx, ti, buf,
                  [&](int i) {
                    const size_t ro = (size_t)i * a.ldb;
                    c_i = at_lane(a.c + ro, boff);
                    x_i = at_lane(a.X + ro, boff);
                    e_i = ld_uniform(a.E + i);
                    uh = T(Ut[i - m0][lane]);
                    vh = T(Vt[i - m0][lane]);
                  },
                  [&](int i, const T (&v)[2]) {
                    if (!act) return;
                    if (one) {  // this row's bf16 right-hand side, to the wave's block below
                      const T rv = at_b_rv(a, i, lam, c_i, x_i, e_i, v[0] + uh, v[1] + vh, d2, cx);
                      rsh[w][i - rbase][lane] = __builtin_bit_cast(unsigned short, (__bf16)float(rv));
                    } else {
                      at_b_row(a, i, inst, lam, c_i, x_i, e_i, v[0] + uh, v[1] + vh, d2, cx);
                    }
                  },
                  &pre);
  if (one && act) {  // rows rbase .. rbase + 7 are one 16-byte block of Rcat per instance
    unsigned b[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) b[r] = rbase + r < a.n ? rsh[w][r][lane] : 0u;  // rows past n: GEMM padding, 0
    *reinterpret_cast<uint4*>(&rcat_at(a.Rcat, a.ldb, rbase, inst)) =
        uint4{b[0] | (b[1] << 16), b[2] | (b[3] << 16), b[4] | (b[5] << 16), b[6] | (b[7] << 16)};
  }
  at_b_partials(a, ti, inst, act, d2, cx, buf);
}

// Phase kernels work on chunks of RPC rows, RW rows per wave: every wave
// issues the loads of all its rows before it computes or stores anything.
constexpr int RPC = 64;
constexpr int RW = RPC / NWV;
constexpr int kPlanes = 8;  // loads in flight per dense row (planes) or light row (entries) in phase C
constexpr int kDenseK = 3;  // phase C: a wave's rows k < kDenseK load their planes with the other rows
constexpr int kNarrowCols = 128;  // batches of at most this many 64-instance blocks run the latency-first kernels

// Phase A over the m-row items of phase C (item blockIdx.y of `items`):
// W (next_w) and the item's partial |(S - b + AX) ./ D|^2 and b'Y per
// instance.  Inside the loop phase C produces these for the rows it
// finishes; this kernel runs only after the layout or the iterate changed
// outside the loop (start, judge, compaction).
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_a(FamArgs<T> a, const int4* __restrict__ items) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const bool act = inst < a.count && !a.done[inst];
  const int4 it = items[blockIdx.y];
  const int r0 = it.x, r1 = it.y;
  double p2 = 0.0, by = 0.0;
  if (act) {
    T sv[RW], yv[RW], av[RW], bv[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = r0 + w + k * NWV;
      if (j < r1) {
        const size_t o = (size_t)j * a.ldb + inst;
        sv[k] = a.S[o]; yv[k] = a.Y[o]; av[k] = a.AX[o]; bv[k] = bval(a, j, inst);
      }
    }
    const T lam = T(a.lam[inst]);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = r0 + w + k * NWV;
      if (j < r1) {
        const T r = ((sv[k] - bv[k]) + av[k]) / T(a.D[j]);
        p2 += double(r) * double(r);
        by += double(bv[k]) * double(yv[k]);
        a.W[(size_t)j * a.ldb + inst] = next_w(bv[k], av[k], sv[k], yv[k], lam);
      }
    }
  }
  waves_sum2(p2, by, buf);
  if (act && w == 0) {
    double* pa = a.pa + ((size_t)blockIdx.y * a.ldb + inst) * 2;
    pa[0] = p2;
    pa[1] = by;
  }
}

// Phase B (one row chunk of the n rows): partial |(c + U) ./ E|^2 and c'X,
// and the x-update right-hand side of every active instance (V = A'W):
//   fp64: R = x - lam c + A'(b - s - lam y), then x+ = G R;
//   mixed: the correction D = R - (I + A'A) x = A'(b - s - lam y - A x) - lam c,
//          then x+ = x + G D, so the bf16 error of the product shrinks with D.
template <typename T, bool MIXED>
__global__ __launch_bounds__(FT) void k_fam_b(FamArgs<T> a) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const bool act = inst < a.count && !a.done[inst];
  const int r0 = blockIdx.y * RPC, r1 = min(a.n, r0 + RPC);
  double d2 = 0.0, cx = 0.0;
  if (act) {
    T cv[RW], xv[RW], uv[RW], vv[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int i = r0 + w + k * NWV;
      if (i < r1) {
        const size_t o = (size_t)i * a.ldb + inst;
        cv[k] = a.c[o]; xv[k] = a.X[o]; uv[k] = a.U[o];
        if (!a.judge) vv[k] = a.V[o];
      }
    }
    // The K planes, added in plane order per row; each plane's loads for all
    // RW rows are issued together (and the unrolled planes ahead of their
    // adds), not one dependent load per row and plane.
#pragma unroll 4
    for (int z = 1; z < a.nsplit; ++z) {
      T pu[RW], pv[RW];
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        const int i = r0 + w + k * NWV;
        if (i < r1) {
          const size_t o = (size_t)i * a.ldb + inst + (size_t)z * a.plane;
          pu[k] = a.U[o];
          if (!a.judge) pv[k] = a.V[o];
        }
      }
#pragma unroll
      for (int k = 0; k < RW; ++k) {
        if (r0 + w + k * NWV < r1) {
          uv[k] += pu[k];
          if (!a.judge) vv[k] += pv[k];
        }
      }
    }
    const T lam = T(a.lam[inst]);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int i = r0 + w + k * NWV;
      if (i < r1) {
        const T r = (cv[k] + uv[k]) / T(a.E[i]);
        d2 += double(r) * double(r);
        cx += double(cv[k]) * double(xv[k]);
        if (a.judge) continue;
        const T rv = MIXED ? vv[k] - lam * cv[k] : xv[k] - lam * cv[k] + vv[k];
        if (MIXED) {
          const __bf16 hi = (__bf16)float(rv);
          rcat_at(a.Rcat, a.ldb, i, inst) = hi;
          if (a.nprod > 1) rcat_at(a.Rcat, a.ldb, a.np + i, inst) = (__bf16)(float(rv) - float(hi));
          if (a.nprod > 2) rcat_at(a.Rcat, a.ldb, 2 * a.np + i, inst) = hi;
        } else {
          a.R[(size_t)i * a.ldb + inst] = rv;
        }
      }
    }
  }
  waves_sum2(d2, cx, buf);
  if (act && w == 0) {
    double* pb = a.pb + ((size_t)blockIdx.y * a.ldb + inst) * 2;
    pb[0] = d2;
    pb[1] = cx;
  }
}

// The stop test of every active instance (pdos.c:95-107) from the chunk
// partials.  A block holds 64 instances; wave w sums chunks w, w + NWV, ...
// of both partial sets (its loads all in flight), and the waves' sums are
// added in wave order, so the totals do not depend on timing.  Wave 0 then
// decides.  In judge mode it compares a warm start's scaled residuals with
// the cold start's and resets rejected instances to zero instead.
// Wave w's share of a stop-test partial table (double2 per chunk and
// instance): chunks w, w + NW, ... < nc summed in that order, with the loads
// of 8 chunks issued before any is used.
template <int NW>
__device__ __forceinline__ void wave_parts(const double* p, int ldb, int inst, int w, int nc, double& s0,
                                           double& s1) {
  for (int c0 = w; c0 < nc; c0 += 8 * NW) {
    double2 q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int c = min(c0 + k * NW, nc - 1);
      q[k] = *reinterpret_cast<const double2*>(p + ((size_t)c * ldb + inst) * 2);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (c0 + k * NW < nc) {
        s0 += q[k].x;
        s1 += q[k].y;
      }
  }
}

// The stop test of column block col (one workgroup of NW waves; `part` is
// NW x LANES double4 of LDS).
template <typename T, int NW = NWV>
__device__ __forceinline__ void decide_block(const FamArgs<T>& a, int col, int chunks_m, int chunks_n,
                                             double4 (*part)[LANES]) {
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = col * LANES + lane;
  // every operand is loaded for every lane (inst < ldb: in bounds) together
  // with the done flag, and finished lanes' sums are dropped after
  const int dn = a.done[inst];
  double4 v = {0.0, 0.0, 0.0, 0.0};  // |pres|^2, b'y, |dres|^2, c'x
  double ps = 1.0, ds = 1.0;
  int it = 0;
  if (w == 0) {
    ps = a.pscale[inst];
    ds = a.dscale[inst];
    it = a.iters[inst];
  }
  wave_parts<NW>(a.pa, a.ldb, inst, w, chunks_m, v.x, v.y);
  wave_parts<NW>(a.pb, a.ldb, inst, w, chunks_n, v.z, v.w);
  const bool act = inst < a.count && !dn;
  if (!act) v = double4{0.0, 0.0, 0.0, 0.0};
  part[w][lane] = v;
  __syncthreads();
  if (w != 0 || !act) return;
  double p2 = 0.0, by = 0.0, d2 = 0.0, cx = 0.0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const double4 q = part[k][lane];
    p2 += q.x;
    by += q.y;
    d2 += q.z;
    cx += q.w;
  }
  const double pres = sqrt(p2), dres = sqrt(d2), pobj = cx, dobj = -by;
  const double eta = fabs(pobj - dobj);
  double* rd = a.red + (size_t)inst * 8;
  if (a.judge) {
    const double mw = fmax(fmax(pres / ps, dres / ds), eta / (1.0 + fabs(pobj) + fabs(dobj)));
    const double mc = fmax((ps - 1.0) / ps, (ds - 1.0) / ds);
    if (!(mw <= mc)) {  // rejected warm start: cold start for this instance
      for (int i = 0; i < a.n; ++i) a.X[(size_t)i * a.ldb + inst] = T(0);
      for (int j = 0; j < a.m; ++j) {
        const size_t o = (size_t)j * a.ldb + inst;
        a.S[o] = T(0); a.Y[o] = T(0); a.AX[o] = T(0);
      }
      a.fallback[inst] = 1;
    }
    return;
  }
  if (it >= a.max_iters) {  // Indeterminate; the residuals of the last check stay
    a.done[inst] = 1;
    return;
  }
  rd[2] = pres; rd[3] = dres; rd[4] = pobj; rd[5] = dobj; rd[6] = eta;
  if (pres < a.eps * ps && dres < a.eps * ds && eta < a.eps * (1.0 + fabs(pobj) + fabs(dobj))) {
    a.done[inst] = 1;
    a.status[inst] = 0;
  } else {
    a.iters[inst] = it + 1;
  }
}


template <typename T>
__global__ __launch_bounds__(FT) void k_fam_decide(FamArgs<T> a, int chunks_m, int chunks_n) {
  __shared__ double4 part[NWV][LANES];
  decide_block(a, blockIdx.x, chunks_m, chunks_n, part);
}

// The A product (as k_fam_spmm<T, 1>: AX = A XN) with the stop test of the
// iteration in the same launch: workgroups 0 .. cols-1 run decide_block for
// their column block, the others the product.  Neither reads what the other
// writes (the product reads XN, the stop test the partials of phase C and the
// A' product), so the iteration saves a kernel boundary.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_ax_decide(FamArgs<T> a, SpmmIn<T, 1> in, int chunks_m, int chunks_n) {
  __shared__ double4 part[NWV][LANES];
  if ((int)blockIdx.x < in.cols) {
    decide_block(a, blockIdx.x, chunks_m, chunks_n, part);
    return;
  }
  double* buf = reinterpret_cast<double*>(&part[0][0]);
  int bx, by;
  spmm_block(in.cols, in.nitems, bx, by, (int)blockIdx.x - in.cols);
  const unsigned boff = (unsigned)(bx * LANES + (threadIdx.x & 63)) * (unsigned)sizeof(T);
  T* Z = a.AX;
  spmm_core<T, 1>(in, bx, by, buf, [](int) {}, [&](int row, const T (&v)[1]) {
    at_lane(Z + (size_t)row * in.ldb, boff) = v[0];
  });
}

// Mixed mode, A with dense rows: phase C forms AX_j for a wave's rows j =
// it.x + w + NWV k of c1 item `item` from the rows' records (rec8): the
// x-update's planes for a dense row, the XN value times the entry for a row
// with one light entry, else the light CSR.  ax_issue puts the loads of all
// rows in flight (the rows' other operands are issued before it);
// ax_finish sums them in a fixed order.
template <typename T>
struct AxLoads {
  int2 rc[RW];
  T av[RW];
  T pv[kDenseK][kPlanes];  // planes of the dense rows among the wave's first kDenseK rows
};

// The wave's records (scalar loads; the table holds one for every (item,
// wave, k), rows past the item's end included).
template <typename T>
__device__ __forceinline__ void ax_records(const FamArgs<T>& a, int item, int w, AxLoads<T>& L) {
#pragma unroll
  for (int k = 0; k < RW; ++k) L.rc[k] = ld_uniform(a.rec8 + ((size_t)item * NWV + w) * RW + k);
}

template <typename T>
__device__ __forceinline__ void ax_issue(const FamArgs<T>& a, int item, int w, int4 it, int inst, AxLoads<T>& L,
                                         bool have_records = false) {
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int j = it.x + w + k * NWV;
    if (j >= it.y) continue;
    if (!have_records) L.rc[k] = ld_uniform(a.rec8 + ((size_t)item * NWV + w) * RW + k);
    L.av[k] = L.rc[k].x >= 0 ? a.XN[(size_t)L.rc[k].x * a.ldb + inst] : T(0);
    if (k < kDenseK && L.rc[k].x == -1) {
#pragma unroll
      for (int u = 0; u < kPlanes; ++u)
        if (u < a.mt) L.pv[k < kDenseK ? k : 0][u] = T(a.AXp[((size_t)u * 32 + L.rc[k].y) * a.ldb + inst]);
    }
  }
}

template <typename T>
__device__ __forceinline__ void ax_finish(const FamArgs<T>& a, int w, int4 it, int inst, const AxLoads<T>& L,
                                          T (&av)[RW]) {
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int j = it.x + w + k * NWV;
    if (j >= it.y) continue;
    T v;
    if (L.rc[k].x >= 0) {
      v = T(__int_as_float(L.rc[k].y)) * L.av[k];
    } else if (L.rc[k].x == -1) {  // dense row: the x-update's planes, in order
      v = T(0);
      int mi = 0;
      if (k < kDenseK) {
#pragma unroll
        for (int u = 0; u < kPlanes; ++u)
          if (u < a.mt) v += L.pv[k < kDenseK ? k : 0][u];
        mi = kPlanes;
      }
      for (; mi < a.mt; ++mi) v += T(a.AXp[((size_t)mi * 32 + L.rc[k].y) * a.ldb + inst]);
    } else {  // several light entries: kPlanes gathers in flight, summed in entry order
      v = T(0);
      const int q1 = ld_uniform(a.lrp + j + 1);
      for (int q0 = ld_uniform(a.lrp + j); q0 < q1; q0 += kPlanes) {
        T g[kPlanes];
        float e[kPlanes];
#pragma unroll
        for (int u = 0; u < kPlanes; ++u) {
          const int q = min(q0 + u, q1 - 1);
          e[u] = ld_uniform(a.lrx + q);
          g[u] = a.XN[(size_t)ld_uniform(a.lri + q) * a.ldb + inst];
        }
#pragma unroll
        for (int u = 0; u < kPlanes; ++u)
          if (q0 + u < q1) v += T(e[u]) * g[u];
      }
    }
    av[k] = v;
  }
}

// Phase C, first pass.  Work items (one per blockIdx.y, host-built, at most
// RPC rows each):
//   (i0, i1, 0, -)  rows of x of finished instances: XN = X (the host swaps
//                   X and XN after every iteration, so active instances take
//                   the GEMM's XN as their new x without a copy)
//   (j0, j1, 1, -)  zero / nonneg rows: s~ = b - AX, relax, project, dual
//                   update; then, for the next iteration, W (next_w) and the
//                   item's stop-test partials (as phase A)
//   (j0, j1, 2, g)  a chunk of second-order cone rows (inside one cone):
//                   s~ and z = s~ - lam y stored, partial |z|^2 (without the
//                   cone's first row) into pc[g]; the cone's first row keeps t0
// With dense rows in A (mixed) AX is formed here and not stored (nothing
// later reads it: compaction moves W and the stop-test partials instead),
// and cone rows store neither s~ nor z: the second pass forms them again
// from the same operands, which costs fewer bytes than storing and reloading.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_c1(FamArgs<T> a, const int4* __restrict__ items) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const bool act = inst < a.count && !a.done[inst];
  const int4 it = items[blockIdx.y];
  if (it.z == 0) {
    if (act || inst >= a.count) return;
    T xv[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int i = it.x + w + k * NWV;
      if (i < it.y) xv[k] = a.X[(size_t)i * a.ldb + inst];
    }
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int i = it.x + w + k * NWV;
      if (i < it.y) a.XN[(size_t)i * a.ldb + inst] = xv[k];
    }
    return;
  }
  const T ra = T(a.relax), rb = T(1.0) - T(a.relax);
  const int cone = it.w, cone_r0 = it.z == 2 ? a.soc_ptr[cone] : 0;
  double n2 = 0.0, p2 = 0.0, by = 0.0;
  const bool hyb = kCorr<T> && a.rec8;
  if (act) {
    T bv[RW], av[RW], sv[RW], yv[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = it.x + w + k * NWV;
      if (j < it.y) {
        const size_t o = (size_t)j * a.ldb + inst;
        bv[k] = bval(a, j, inst); sv[k] = a.S[o]; yv[k] = a.Y[o];
        if (!hyb) av[k] = a.AX[o];
      }
    }
    if constexpr (kCorr<T>) {
      if (hyb) {
        AxLoads<T> L;
        ax_issue(a, blockIdx.y, w, it, inst, L);
        ax_finish(a, w, it, inst, L, av);
      }
    }
    const T lam = T(a.lam[inst]);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = it.x + w + k * NWV;
      if (j < it.y) {
        const size_t o = (size_t)j * a.ldb + inst;
        T st = bv[k] - av[k];
        st = ra * st + rb * sv[k];
        const T z = st - lam * yv[k];
        if (it.z == 1) {
          const T s = j < a.f ? T(0) : (z < T(0) ? T(0) : z);
          const T y = yv[k] + (s - st) / lam;
          a.S[o] = s;
          a.Y[o] = y;
          const T r = ((s - bv[k]) + av[k]) / T(a.D[j]);
          p2 += double(r) * double(r);
          by += double(bv[k]) * double(y);
          a.W[o] = next_w(bv[k], av[k], s, y, lam);
        } else {
          if (!hyb) {
            a.W[o] = st;  // s~ for the dual update
            a.S[o] = z;
          }
          if (j > cone_r0) n2 += double(z) * double(z);
          else a.pt0[(size_t)cone * a.ldb + inst] = double(z);  // t0, read by every chunk of the cone
        }
      }
    }
  }
  if (it.z == 1) {
    waves_sum2(p2, by, buf);
    if (act && w == 0) {
      double* pa = a.pa + ((size_t)(blockIdx.y - a.nx_items) * a.ldb + inst) * 2;
      pa[0] = p2;
      pa[1] = by;
    }
    return;
  }
  n2 = waves_sum(n2, buf);
  if (act && w == 0) a.pc[(size_t)blockIdx.y * a.ldb + inst] = n2;
}

// Phase C, second pass over the cone chunks: the cone's norm is the sum of
// its chunks' partials (fixed order) and t0 comes from the first pass, so no
// block depends on another block's writes; project (cones.c:26-47), dual
// update; then W (next_w) and the chunk's stop-test partials for the next
// iteration.  Item (j0, j1, first chunk of the cone, chunks in the cone);
// cone[item] is the cone index.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_c2(FamArgs<T> a, const int4* __restrict__ items,
                                               const int* __restrict__ cone) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const bool act = inst < a.count && !a.done[inst];
  const int4 it = items[blockIdx.y];
  const int cn = cone[blockIdx.y];
  // mixed mode without relaxation: s~ = b - AX exactly, so the residual and
  // W come from s~ and AX, b are not read again
  const bool from_st = kCorr<T> && a.relax == 1.0;
  // A with dense rows (mixed): s~ and z are formed again as in the first pass
  const bool hyb = kCorr<T> && a.rec8;
  double p2 = 0.0, by = 0.0;
  if (act) {
    const T lam = T(a.lam[inst]);
    const T t0 = T(a.pt0[(size_t)cn * a.ldb + inst]);
    T zv[RW], wv[RW], yv[RW], av[RW], bv[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = it.x + w + k * NWV;
      if (j < it.y) {
        const size_t o = (size_t)j * a.ldb + inst;
        zv[k] = a.S[o]; yv[k] = a.Y[o];
        bv[k] = bval(a, j, inst);
        if (!hyb) {
          wv[k] = a.W[o];
          if (!from_st) av[k] = a.AX[o];
        }
      }
    }
    if constexpr (kCorr<T>) {
      if (hyb) {  // zv holds S of the last iteration here
        AxLoads<T> L;
        ax_issue(a, a.nx_items + a.nlin_items + blockIdx.y, w, it, inst, L);
        ax_finish(a, w, it, inst, L, av);
        const T ra = T(a.relax), rb = T(1.0) - T(a.relax);
#pragma unroll
        for (int k = 0; k < RW; ++k) {
          if (it.x + w + k * NWV >= it.y) continue;
          T st = bv[k] - av[k];
          st = ra * st + rb * zv[k];
          wv[k] = st;
          zv[k] = st - lam * yv[k];
        }
      }
    }
    // the cone's chunk partials in order, 16 loads issued at a time
    double n2 = 0.0;
    const int gend = it.z + it.w;
    for (int g0 = it.z; g0 < gend; g0 += 16) {
      double pv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) pv[u] = a.pc[(size_t)min(g0 + u, gend - 1) * a.ldb + inst];
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (g0 + u < gend) n2 += pv[u];
    }
    const int r0 = a.soc_ptr[cn];
    const T ns = T(sqrt(n2));
    const int mode = ns <= t0 ? 0 : (ns <= -t0 ? 1 : 2);
    const T al = (ns + t0) / T(2);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = it.x + w + k * NWV;
      if (j < it.y) {
        const size_t o = (size_t)j * a.ldb + inst;
        const T z = zv[k];
        const T sv = mode == 0 ? z : (mode == 1 ? T(0) : (j == r0 ? al : al * z / ns));
        const T y = yv[k] + (sv - wv[k]) / lam;
        a.Y[o] = y;
        a.S[o] = sv;
        const T r = from_st ? (sv - wv[k]) / T(a.D[j]) : ((sv - bv[k]) + av[k]) / T(a.D[j]);
        p2 += double(r) * double(r);
        by += double(bv[k]) * double(y);
        a.W[o] = from_st ? wv[k] - (sv + lam * y) : next_w(bv[k], av[k], sv, y, lam);
      }
    }
  }
  waves_sum2(p2, by, buf);
  if (act && w == 0) {
    double* pa = a.pa + ((size_t)(a.nlin_items + blockIdx.y) * a.ldb + inst) * 2;
    pa[0] = p2;
    pa[1] = by;
  }
}

// Phase C first pass for narrow batches (A with dense rows, mixed): as
// k_fam_c1 with its records read first as one batch of scalar loads and
// every operand as one batch of vector loads for every lane (in bounds for
// inst < ldb); only the stores wait for `done`.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_c1_ll(FamArgs<T> a, const int4* __restrict__ items) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const int4 it = ld_uniform(items + blockIdx.y);
  AxLoads<T> L;
  ax_records(a, blockIdx.y, w, L);
  const int dn = a.done[inst];
  const T lam = T(a.lam[inst]);
  if (it.z == 0) {  // x rows of finished instances: XN = X
    T xv[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int i = it.x + w + k * NWV;
      if (i < it.y) xv[k] = a.X[(size_t)i * a.ldb + inst];
    }
    if (inst >= a.count || !dn) return;
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int i = it.x + w + k * NWV;
      if (i < it.y) a.XN[(size_t)i * a.ldb + inst] = xv[k];
    }
    return;
  }
  const int cone = it.w;
  const int cone_r0 = it.z == 2 ? ld_uniform(a.soc_ptr + cone) : 0;
  T bv[RW], av[RW], sv[RW], yv[RW];
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int j = it.x + w + k * NWV;
    if (j < it.y) {
      const size_t o = (size_t)j * a.ldb + inst;
      bv[k] = bval(a, j, inst); sv[k] = a.S[o]; yv[k] = a.Y[o];
    }
  }
  ax_issue(a, 0, w, it, inst, L, true);
  const bool act = inst < a.count && !dn;
  ax_finish(a, w, it, inst, L, av);
  const T ra = T(a.relax), rb = T(1.0) - T(a.relax);
  double n2 = 0.0, p2 = 0.0, by = 0.0;
  if (act) {
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = it.x + w + k * NWV;
      if (j < it.y) {
        const size_t o = (size_t)j * a.ldb + inst;
        T st = bv[k] - av[k];
        st = ra * st + rb * sv[k];
        const T z = st - lam * yv[k];
        if (it.z == 1) {
          const T s = j < a.f ? T(0) : (z < T(0) ? T(0) : z);
          const T y = yv[k] + (s - st) / lam;
          a.S[o] = s;
          a.Y[o] = y;
          const T r = ((s - bv[k]) + av[k]) / T(a.D[j]);
          p2 += double(r) * double(r);
          by += double(bv[k]) * double(y);
          a.W[o] = next_w(bv[k], av[k], s, y, lam);
        } else {
          if (j > cone_r0) n2 += double(z) * double(z);
          else a.pt0[(size_t)cone * a.ldb + inst] = double(z);  // t0, read by every chunk of the cone
        }
      }
    }
  }
  if (it.z == 1) {
    waves_sum2(p2, by, buf);
    if (act && w == 0) {
      double* pa = a.pa + ((size_t)(blockIdx.y - a.nx_items) * a.ldb + inst) * 2;
      pa[0] = p2;
      pa[1] = by;
    }
    return;
  }
  n2 = waves_sum(n2, buf);
  if (act && w == 0) a.pc[(size_t)blockIdx.y * a.ldb + inst] = n2;
}

// Phase C second pass for narrow batches, where each launch is a chain of
// memory round trips rather than a stream: A with dense rows (mixed) only.
// The item, cone and row records are read first as one batch of scalar
// loads (none depends on data), then every operand of the wave's rows (S, Y,
// b, the XN values and planes of AX, the cone's chunk partials, t0, lambda)
// as one batch of vector loads issued for every lane, finished or not (all
// in bounds for inst < ldb); only the stores wait for `done`.  Same
// arithmetic as k_fam_c2.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_c2_ll(FamArgs<T> a, const int4* __restrict__ items,
                                                  const int* __restrict__ cone) {
  __shared__ double buf[2 * NWV * LANES];
  const int lane = threadIdx.x & 63, w = wave_id();
  const int inst = blockIdx.x * LANES + lane;
  const int4 it = ld_uniform(items + blockIdx.y);
  const int cn = ld_uniform(cone + blockIdx.y);
  AxLoads<T> L;
  ax_records(a, a.nx_items + a.nlin_items + blockIdx.y, w, L);
  const int dn = a.done[inst];
  const T lam = T(a.lam[inst]);
  const int r0 = ld_uniform(a.soc_ptr + cn);
  const T t0 = T(a.pt0[(size_t)cn * a.ldb + inst]);
  T zv[RW], wv[RW], yv[RW], av[RW], bv[RW];
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int j = it.x + w + k * NWV;
    if (j < it.y) {
      const size_t o = (size_t)j * a.ldb + inst;
      zv[k] = a.S[o]; yv[k] = a.Y[o];
      bv[k] = bval(a, j, inst);
    }
  }
  ax_issue(a, 0, w, it, inst, L, true);
  const int gend = it.z + it.w;
  double pv[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) pv[u] = a.pc[(size_t)min(it.z + u, gend - 1) * a.ldb + inst];
  const bool act = inst < a.count && !dn;
  ax_finish(a, w, it, inst, L, av);
  const T ra = T(a.relax), rb = T(1.0) - T(a.relax);
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    if (it.x + w + k * NWV >= it.y) continue;
    T st = bv[k] - av[k];
    st = ra * st + rb * zv[k];
    wv[k] = st;
    zv[k] = st - lam * yv[k];
  }
  double n2 = 0.0;
#pragma unroll
  for (int u = 0; u < 16; ++u)
    if (it.z + u < gend) n2 += pv[u];
  for (int g0 = it.z + 16; g0 < gend; g0 += 16) {  // cones of more than 16 chunks
    double qv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) qv[u] = a.pc[(size_t)min(g0 + u, gend - 1) * a.ldb + inst];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (g0 + u < gend) n2 += qv[u];
  }
  const bool from_st = a.relax == 1.0;
  double p2 = 0.0, by = 0.0;
  if (act) {
    const T ns = T(sqrt(n2));
    const int mode = ns <= t0 ? 0 : (ns <= -t0 ? 1 : 2);
    const T al = (ns + t0) / T(2);
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int j = it.x + w + k * NWV;
      if (j < it.y) {
        const size_t o = (size_t)j * a.ldb + inst;
        const T z = zv[k];
        const T sv = mode == 0 ? z : (mode == 1 ? T(0) : (j == r0 ? al : al * z / ns));
        const T y = yv[k] + (sv - wv[k]) / lam;
        a.Y[o] = y;
        a.S[o] = sv;
        const T r = from_st ? (sv - wv[k]) / T(a.D[j]) : ((sv - bv[k]) + av[k]) / T(a.D[j]);
        p2 += double(r) * double(r);
        by += double(bv[k]) * double(y);
        a.W[o] = from_st ? wv[k] - (sv + lam * y) : next_w(bv[k], av[k], sv, y, lam);
      }
    }
  }
  waves_sum2(p2, by, buf);
  if (act && w == 0) {
    double* pa = a.pa + ((size_t)(a.nlin_items + blockIdx.y) * a.ldb + inst) * 2;
    pa[0] = p2;
    pa[1] = by;
  }
}

// Row-major m x n dense copy with the lower triangle mirrored into the upper.
template <typename T>
__global__ void k_symmetrize_convert(int n, const double* __restrict__ G, T* __restrict__ out, int ld_out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    // column-major lower triangle from potri: element (r, c) with r >= c at c * n + r
    const int r = i > j ? i : j, c = i > j ? j : i;
    out[(size_t)i * ld_out + j] = T(G[(size_t)c * n + r]);
  }
}

// Gcat = [G_hi, G_hi, G_lo] (np x 3np, bf16) from the fp64 symmetric inverse,
// so Gcat [R_hi; R_lo; R_hi] = G_hi R_hi + G_hi R_lo + G_lo R_hi.
__global__ void k_split_g(int n, int np, const double* __restrict__ G, __bf16* __restrict__ Gcat) {
  const long total = (long)np * np;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / np), j = (int)(e % np);
    float v = 0.f;
    if (i < n && j < n) {
      const int r = i > j ? i : j, c = i > j ? j : i;
      v = float(G[(size_t)c * n + r]);
    }
    const __bf16 hi = (__bf16)v;
    const __bf16 lo = (__bf16)(v - float(hi));
    const size_t row = (size_t)i * 3 * np;  // pairs with [R_hi; R_lo; R_hi]
    Gcat[row + j] = hi;
    Gcat[row + np + j] = hi;
    Gcat[row + 2 * np + j] = lo;
  }
}

typedef __bf16 xbf16x8 __attribute__((ext_vector_type(8)));
typedef float xf32x4 __attribute__((ext_vector_type(4)));

// MFMA 16x16x32 K steps of one wave and batch of the x-update: all their
// operand loads are issued before the first product.
constexpr int kXuSteps = 4;

// The mixed-precision x-update XN = X + Gcat[:, :K] Rcat (K = nprod np):
// Gcat row-major np x ldg bf16, Rcat in 8-row blocks (rcat_at), X and XN
// row-major np x ldb binary32.  A workgroup of 4 waves computes a 64 x 64
// tile (rows m0.., instances n0..): wave q multiplies the K steps of 32
// q, q + 4, ... for the whole tile, loading its operands straight from
// memory (A: 16 contiguous bytes of a row of G; B: one 16-byte Rcat block),
// so a tile costs one round trip per kXuSteps steps of a wave instead of a
// pass through LDS per step.  The four partial tiles meet in LDS and are
// summed in wave order (deterministic) with the addend X.  Grid 1-D over
// (row tile, column tile); with ntile % 8 == 0 the row tiles of one column
// tile go to one XCD, consecutively, so their Rcat reads share its L2.  For an
// A with dense rows the launch also carries the iteration's stop test
// (workgroups past the tiles, one per column block).
__global__ __launch_bounds__(256) void k_fam_xupdate(int mt, int nt, int ksteps, const __bf16* __restrict__ G,
                                                     int ldg, const __bf16* __restrict__ R, int ldb,
                                                     const float* __restrict__ X, float* __restrict__ XN,
                                                     const float* __restrict__ Ahr, float* __restrict__ AXp,
                                                     FamArgs<float> fa, int chunks_m, int chunks_n) {
  __shared__ float red[4][64][65];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int L = blockIdx.x;
  if (L >= mt * nt) {  // the iteration's stop test for column block L - mt nt (A with dense rows)
    decide_block<float, 4>(fa, L - mt * nt, chunks_m, chunks_n, reinterpret_cast<double4(*)[LANES]>(&red[0][0][0]));
    return;
  }
  int mi, ni;
  if ((nt & 7) == 0) {
    const int seq = L >> 3;
    ni = (seq / mt) * 8 + (L & 7);
    mi = seq % mt;
  } else {
    ni = L / mt;
    mi = L % mt;
  }
  const int m0 = mi * 64, n0 = ni * 64;
  // A with dense rows: this tile's values of A_h for the planes below, in
  // flight with the product's operands
  float ahv[4][2];
  if (AXp) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int rg = 0; rg < 2; ++rg) ahv[s][rg] = Ahr[((size_t)((m0 + 16 * q) / 4 + s) * 2 + rg) * 64 + lane];
  }
  // addend rows 16 q .. 16 q + 15 of the tile, column `lane`, for the epilogue
  float xa[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) xa[r] = X[(size_t)(m0 + 16 * q + r) * ldb + n0 + lane];
  xf32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = xf32x4{0.f, 0.f, 0.f, 0.f};
  const int fr = lane & 15, fk = lane >> 4;  // operand row / column, 8-wide K group
  for (int s0 = q; s0 < ksteps; s0 += 4 * kXuSteps) {
    xbf16x8 af[kXuSteps][4], bv[kXuSteps][4];
#pragma unroll
    for (int t = 0; t < kXuSteps; ++t) {
      const int st = min(s0 + 4 * t, ksteps - 1);  // past the end: repeat the last step, not used
      const int k = st * 32 + fk * 8;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        af[t][i] = *reinterpret_cast<const xbf16x8*>(G + (size_t)(m0 + i * 16 + fr) * ldg + k);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bv[t][j] = *reinterpret_cast<const xbf16x8*>(R + ((size_t)(k >> 3) * ldb + n0 + j * 16 + fr) * 8);
    }
#pragma unroll
    for (int t = 0; t < kXuSteps; ++t) {
      if (s0 + 4 * t < ksteps) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[t][i], bv[t][j], acc[i][j], 0, 0, 0);
      }
    }
  }
  // accumulator lane l: rows (l / 16) * 4 + r, column l % 16 of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[q][i * 16 + fk * 4 + r][j * 16 + fr] = acc[i][j][r];
  __syncthreads();
  float xn[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = 16 * q + r;
    const float v = ((red[0][row][lane] + red[1][row][lane]) + red[2][row][lane]) + red[3][row][lane];
    xn[r] = xa[r] + v;
    XN[(size_t)(m0 + row) * ldb + n0 + lane] = xn[r];
  }
  if (!AXp) return;
  // A with dense rows: this tile's part of AX_h = A_h XN, the 32 dense rows
  // over K rows m0 .. m0 + 63, on the fp32 MFMA 16x16x4 (wave q: K rows
  // m0 + 16 q ..); the four wave partials meet in LDS and each wave adds 8
  // rows of them in wave order into plane mi of AXp (32 x ldb per row tile;
  // phase C sums the planes).  Each wave rewrites only the rows of
  // red[0..3] it read above, so the XN tile can go to red[0] at once.
#pragma unroll
  for (int r = 0; r < 16; ++r) red[0][16 * q + r][lane] = xn[r];
  __syncthreads();
  hf32x4 ah[2][4];
#pragma unroll
  for (int rg = 0; rg < 2; ++rg)
#pragma unroll
    for (int cg = 0; cg < 4; ++cg) ah[rg][cg] = hf32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s) {  // K quad (m0 + 16 q) / 4 + s
#pragma unroll
    for (int cg = 0; cg < 4; ++cg) {
      const float xb = red[0][16 * q + 4 * s + fk][cg * 16 + fr];
#pragma unroll
      for (int rg = 0; rg < 2; ++rg)
        ah[rg][cg] = __builtin_amdgcn_mfma_f32_16x16x4f32(ahv[s][rg], xb, ah[rg][cg], 0, 0, 0);
    }
  }
  // partial of wave q: rows (q & 1) * 32 .. of red[1 + q / 2]
  float(*pq)[65] = &red[1 + (q >> 1)][(q & 1) * 32];
#pragma unroll
  for (int rg = 0; rg < 2; ++rg)
#pragma unroll
    for (int cg = 0; cg < 4; ++cg)
#pragma unroll
      for (int r = 0; r < 4; ++r) pq[rg * 16 + fk * 4 + r][cg * 16 + fr] = ah[rg][cg][r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int row = 8 * q + r;
    const float v = ((red[1][row][lane] + red[1][32 + row][lane]) + red[2][row][lane]) + red[2][32 + row][lane];
    AXp[((size_t)mi * 32 + row) * ldb + n0 + lane] = v;
  }
}


// Unscaled outputs, instance-major binary64 (pdos.c:251-294): out[inst][r] =
// Z[r][inst] * scale[r] (x = E x, y = D y) or / scale[r] (s = s / D), one
// rounding each, as the host's unscale.
template <typename T>
__global__ void k_gather_out(int rows, int count, int ldb, const T* __restrict__ Z, const double* __restrict__ scale,
                             int divide, double* __restrict__ out) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)rows * count; e += (long)gridDim.x * blockDim.x) {
    const int inst = (int)(e / rows), r = (int)(e % rows);
    const double z = double(Z[(size_t)r * ldb + inst]);
    out[e] = divide ? z / scale[r] : scale[r] * z;
  }
}

// Per-instance setup of a loaded batch on the device (common.h:69-215 for
// the family's per-instance data): thread t scales instance t's c by E0 (the
// column scaling before the x1e3) and writes it feature-major with the x1e3,
// the same for b by D when b is per instance, and forms the stop-test scales
// 1 + |b|, 1 + |c| and lambda = (1e-15 + |b~|) / (1e-15 + |c~|).  Each norm is
// a sequential sum in index order with separately rounded products and sums
// (contraction off: no fused multiply-add), the host's arithmetic bit for
// bit.  With b shared, ps_sh is its 1 + |b| and nbs_sh its |b~|^2.
// Padding columns get c = b = 0 and scales 1.
template <typename T>
__global__ void k_fam_scale_in(int cnt, int n, int m, int ldb, const double* __restrict__ C,
                               const double* __restrict__ B, const double* __restrict__ E0,
                               const double* __restrict__ D, double ps_sh, double nbs_sh, T* __restrict__ cf,
                               T* __restrict__ bf, double* __restrict__ lam, double* __restrict__ ps,
                               double* __restrict__ ds) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ldb) return;
  if (t >= cnt) {
    for (int i = 0; i < n; ++i) cf[(size_t)i * ldb + t] = T(0);
    if (bf)
      for (int j = 0; j < m; ++j) bf[(size_t)j * ldb + t] = T(0);
    lam[t] = ps[t] = ds[t] = 1.0;
    return;
  }
  double nc = 0.0, ncs = 0.0, nb = 0.0, nbs = nbs_sh;
  const double* c = C + (size_t)t * n;
#pragma unroll 8
  for (int i = 0; i < n; ++i) {
    const double ci = c[i];
    nc += ci * ci;
    const double cs = ci * E0[i];
    ncs += cs * cs;
    cf[(size_t)i * ldb + t] = T(cs * 1e3);
  }
  if (B) {
    nbs = 0.0;
    const double* b = B + (size_t)t * m;
#pragma unroll 8
    for (int j = 0; j < m; ++j) {
      const double bj = b[j];
      nb += bj * bj;
      const double bs = bj * D[j];
      nbs += bs * bs;
      bf[(size_t)j * ldb + t] = T(bs);
    }
  }
  ps[t] = B ? 1.0 + sqrt(nb) : ps_sh;
  ds[t] = 1.0 + sqrt(nc);
  lam[t] = (1e-15 + sqrt(nbs)) / (1e-15 + sqrt(ncs));
}

// Column moves (compaction, output scatter): a lane copies one column's
// rows from src (row stride sld, column sc) to dst (row stride dld, column
// dc).  The rows are dealt over the block's waves and over gridDim.y blocks
// (kMoveSplit), and each wave issues kMoveU rows' loads before their stores:
// src and dst never overlap, but the compiler cannot know that, and one load
// in flight per wave measured ~130 GB/s.
constexpr int kMoveSplit = 4, kMoveU = 4;
template <typename T>
__device__ __forceinline__ void move_rows(const T* src, size_t sld, int sc, T* dst, size_t dld, int dc, int rows,
                                          bool live, int w) {
  const int step = NWV * gridDim.y;
  for (int r0 = w + NWV * blockIdx.y; r0 < rows; r0 += step * kMoveU) {
    T v[kMoveU];
#pragma unroll
    for (int u = 0; u < kMoveU; ++u) {
      const int r = r0 + u * step;
      v[u] = (live && r < rows) ? src[(size_t)r * sld + sc] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kMoveU; ++u) {
      const int r = r0 + u * step;
      if (r < rows) dst[(size_t)r * dld + dc] = v[u];
    }
  }
}

// Writes the finished columns of a working set to the output store
// (feature-major, ld = ldo, indexed by the original instance).
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_scatter(FamArgs<T> a, int ldo, T* __restrict__ Xo, T* __restrict__ Yo,
                                                    T* __restrict__ So, int* __restrict__ Io,
                                                    double* __restrict__ Ro) {
  const int lane = threadIdx.x & 63, w = wave_id();
  const int col = blockIdx.x * LANES + lane;
  if (col >= a.count || !a.done[col]) return;
  const int t = a.orig[col];
  move_rows(a.X, a.ldb, col, Xo, ldo, t, a.n, true, w);
  move_rows(a.Y, a.ldb, col, Yo, ldo, t, a.m, true, w);
  move_rows(a.S, a.ldb, col, So, ldo, t, a.m, true, w);
  if (w == 0 && blockIdx.y == 0) {
    Io[t] = a.iters[col];
    Io[ldo + t] = a.status[col];
    Io[2 * ldo + t] = a.fallback[col];
    for (int k = 0; k < 8; ++k) Ro[(size_t)k * ldo + t] = a.red[(size_t)col * 8 + k];
  }
}

// Gathers the columns `list` of working set s into the dense prefix of d
// (d.count live columns, d.ldb - d.count padding columns left inactive).
// With pa_items > 0 (A with dense rows, mixed: phase C does not keep AX) W
// and the pa_items stop-test partials move instead of AX, so the moved state
// is complete without a phase A pass.
template <typename T>
__global__ __launch_bounds__(FT) void k_fam_move(FamArgs<T> s, FamArgs<T> d, const int* __restrict__ list,
                                                 int pa_items) {
  const int lane = threadIdx.x & 63, w = wave_id();
  const int col = blockIdx.x * LANES + lane;
  const bool live = col < d.count;
  const int sc = live ? list[col] : 0;
  auto mv = [&](const T* src, T* dst, int rows) { move_rows(src, s.ldb, sc, dst, d.ldb, col, rows, live, w); };
  const int r0 = w + NWV * blockIdx.y, rstep = NWV * gridDim.y;
  mv(s.X, d.X, d.n);
  mv(s.S, d.S, d.m);
  mv(s.Y, d.Y, d.m);
  if (pa_items > 0) {
    mv(s.W, d.W, d.m);
    for (int it = r0; it < pa_items; it += rstep) {
      const double2 v = live ? *reinterpret_cast<const double2*>(s.pa + ((size_t)it * s.ldb + sc) * 2)
                             : double2{0.0, 0.0};
      *reinterpret_cast<double2*>(d.pa + ((size_t)it * d.ldb + col) * 2) = v;
    }
  } else {
    mv(s.AX, d.AX, d.m);
  }
  mv(s.c, d.c, d.n);
  if (d.b_inst) mv(s.b, d.b, d.m);
  for (int r = d.n + r0; r < d.np; r += rstep) {  // padding rows of the GEMM operands
    d.X[(size_t)r * d.ldb + col] = T(0);
    if (d.Rcat)
      for (int k = 0; k < 3; ++k) rcat_at(d.Rcat, d.ldb, k * d.np + r, col) = (__bf16)0.f;
  }
  if (w == 0 && blockIdx.y == 0) {
    d.lam[col] = live ? s.lam[sc] : 1.0;
    d.pscale[col] = live ? s.pscale[sc] : 1.0;
    d.dscale[col] = live ? s.dscale[sc] : 1.0;
    d.done[col] = live ? 0 : 1;
    d.status[col] = live ? s.status[sc] : 1;
    d.iters[col] = live ? s.iters[sc] : 0;
    d.fallback[col] = live ? s.fallback[sc] : 0;
    d.orig[col] = live ? s.orig[sc] : 0;
    for (int k = 0; k < 8; ++k) d.red[(size_t)col * 8 + k] = live ? s.red[(size_t)sc * 8 + k] : -1.0;
  }
}

__global__ void k_iota(int n, int* __restrict__ v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

int pad64(int v) { return (v + 63) & ~63; }

}  // namespace

// ---------------------------------------------------------------------------

struct Family {
  Setup S;              // shared matrix, scalings, cones (b, c of the template data)
  Params par;
  bool fp64 = false;
  int device = 0;
  hipStream_t st = nullptr;
  rocblas_handle blas = nullptr;
  int n = 0, m = 0, np = 0;
  // shared device data
  void *rp = nullptr, *ri = nullptr, *rx = nullptr, *cp = nullptr, *ci = nullptr, *cx = nullptr;
  void *soc = nullptr, *D = nullptr, *E = nullptr, *E0 = nullptr, *bshared = nullptr;
  // the loaded batch's raw inputs (instance-major binary64) and the fetch's
  // instance-major outputs, kept with the batch
  void *inC = nullptr, *inB = nullptr, *gout = nullptr;
  void* Ad = nullptr;      // dense scaled A (m x n row-major, bulk type) when A is dense
  // mixed precision, dense A: both products run as the hand-written split
  // bf16 MFMA kernels (dense_gemm.hip; A = A_hi + A_lo, three products with
  // fp32 accumulation) on A's split copies in MFMA operand order:
  //   ATh, ATl  A' rows, dense_row_pad(n) x m64;  ARh, ARl  A rows,
  //   dense_row_pad(m) x n64 (zero padded)
  //   Upl, Vpl  the A' products' split-K planes
  void *ATh = nullptr, *ATl = nullptr, *ARh = nullptr, *ARl = nullptr, *Upl = nullptr, *Vpl = nullptr;
  int m64 = 0, n64 = 0;
  bool split_dense = false;
  bool dense = false;
  void* Gd = nullptr;      // fp64 G (n x n row-major, symmetric) for dgemm
  void* Gcat = nullptr;    // mixed: bf16 [G_hi, G_lo, G_hi]
  // per-batch device data: the loaded constants, two working sets (the
  // second receives the compacted active columns), and the output store
  int count = 0, ldb = 0, b_inst = 0;
  std::vector<void*> batch_allocs;
  struct Set {
    void *X, *XN, *S, *Y, *AX, *W, *U, *V, *R, *Rcat, *c, *b, *lam, *ps, *ds, *ints, *red, *orig, *pa, *pb, *pc, *pt0;
  } set[2];
  void *base_c = nullptr, *base_b = nullptr, *base_lam = nullptr, *base_ps = nullptr, *base_ds = nullptr;
  void *Xo = nullptr, *Yo = nullptr, *So = nullptr, *Io = nullptr, *Ro = nullptr;
  int* dlist = nullptr;
  int* hflags = nullptr;    // pinned: done flags of the last two chunks (2 x ldb)
  hipEvent_t fev[2] = {nullptr, nullptr};
  // fetch: the outputs come down through two pinned staging chunks (events
  // sev) while the host copies the previous chunk out on several threads
  static constexpr size_t kFetchChunk = (size_t)32 << 20;
  char* stage = nullptr;
  hipEvent_t sev[2] = {nullptr, nullptr};
  int compactions = 0;
  int nprod = 1;          // mixed: bf16 products of the split GEMM (1, 2 or 3)
  // work tables (device), built once per family: SpMM wave tasks for A and
  // A' at two granularities ([0] up to 64 entries per task, [1] up to 16, for
  // small compacted batches) and the phase C items (RPC rows each)
  struct Tasks {
    int4 *A = nullptr, *At = nullptr;
    int nA = 0, nAt = 0;  // blocks (NWV tasks each)
  } tasks[2];
  // mixed precision, A with dense rows: k_fam_at_hyb's operands (nh dense
  // rows; the light CSR of A' and one task per wave and 64-row tile)
  int nh = 0;
  DenseRows drows{};
  float* Ah = nullptr;
  int *li = nullptr, *lrow = nullptr;
  void* lx = nullptr;
  int4* ltasks = nullptr;
  // ... and for the A product, which the x-update (dense rows: its planes
  // AXp) and phase C (the other rows, from XN) form: A_h (MFMA order), the
  // light CSR of A (dense rows empty) and phase C's per-row records (a dense
  // row's record holds its slot among the dense rows)
  float* Ahr = nullptr;
  int *lrpd = nullptr, *lri = nullptr;
  float* lrxf = nullptr;
  int2* rec8d = nullptr;
  float* AXp = nullptr;  // per batch: the x-update's row-tile planes of A_h XN (np / 64 x 32 x ldb)
  bool hybrid() const { return nh > 0 && !fp64 && !dense; }
  const Tasks* curA = &tasks[0];   // task table of the A product for the current width
  const Tasks* curAt = &tasks[0];  // and of the A' product
  int num_cu = 256;
  // Task granularity of a product over `cols` column blocks: the coarse table
  // (64 entries per task: an index load and four gather rounds per wave) or
  // the fine one (16 entries: one of each), whichever needs fewer sequential
  // memory round trips, counted as rounds of resident workgroups (4 of 512
  // threads per CU) times round trips per task.
  int pick_tasks(int cols, int n_coarse, int n_fine) const {
    const long slots = 4L * num_cu;
    const long rc = (cols * (long)n_coarse + slots - 1) / slots, rf = (cols * (long)n_fine + slots - 1) / slots;
    return rc * 5 <= rf * 2 ? 0 : 1;
  }
  int *rowA = nullptr, *rowAt = nullptr;  // row of every CSR entry of A, of A'
  int4 *c1 = nullptr, *c2 = nullptr;
  int* c2_cone = nullptr;
  int nc1 = 0, nc2 = 0, nx_items = 0, nlin_items = 0;  // c1: x items, then zero/nonneg items, then cone chunks
  int m_items() const { return nc1 - nx_items; }       // stop-test partials over the m rows
  bool need_prologue = true;                           // W and the partials must be formed by k_fam_a
  double last_ms = 0;
  // optional per-kernel timing: HIP events around every launch of the loop
  bool timing = false;
  enum { K_A = 0, K_AT, K_B, K_GEMM, K_AX, K_C, K_KINDS };
  double kind_ms[K_KINDS] = {0};
  long kind_calls[K_KINDS] = {0};
  double kind_cols[K_KINDS] = {0};  // columns (instances incl. padding) the launches covered
  int cur_ld = 0;
  std::vector<hipEvent_t> evs;
  std::vector<int> ev_kind, ev_ld;
  void mark(int kind) {  // event before launching `kind` (kind < 0: closing event)
    if (!timing) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, st);
    evs.push_back(e);
    ev_kind.push_back(kind);
    ev_ld.push_back(cur_ld);
  }
  void collect() {
    if (!timing) return;
    (void)hipStreamSynchronize(st);
    for (size_t i = 0; i + 1 < evs.size(); ++i) {
      if (ev_kind[i] < 0) continue;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, evs[i], evs[i + 1]);
      kind_ms[ev_kind[i]] += ms;
      kind_calls[ev_kind[i]] += 1;
      kind_cols[ev_kind[i]] += ev_ld[i];
    }
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    evs.clear();
    ev_kind.clear();
    ev_ld.clear();
  }
  std::vector<void*> allocs;

  ~Family() {
    free_batch();
    for (hipEvent_t& e : fev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : sev)
      if (e) (void)hipEventDestroy(e);
    if (stage) (void)hipHostFree(stage);
    for (void* p : allocs) (void)hipFree(p);
    if (blas) rocblas_destroy_handle(blas);
    if (st) (void)hipStreamDestroy(st);
  }
  void* dalloc(size_t bytes, std::vector<void*>& list, std::string& err) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) { err = "hipMalloc failed"; return nullptr; }
    (void)hipMemsetAsync(p, 0, bytes ? bytes : 1, st);
    list.push_back(p);
    return p;
  }
  template <typename V>
  void* upload(const std::vector<V>& v, std::string& err) {
    void* p = dalloc(v.size() * sizeof(V), allocs, err);
    if (p) (void)hipMemcpyAsync(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice, st);
    return p;
  }
  void* upload_vals(const std::vector<double>& v, std::string& err) {
    if (fp64) return upload(v, err);
    std::vector<float> f(v.begin(), v.end());
    return upload(f, err);
  }
  void free_batch() {
    for (void* p : batch_allocs) (void)hipFree(p);
    batch_allocs.clear();
    if (hflags) (void)hipHostFree(hflags);
    hflags = nullptr;
    count = 0;
  }

  bool create(const Data* d, const Cone* k, const PdosOptions& o, std::string& err);
  bool load(int cnt, const double* C, const double* B, std::string& err);
  bool alloc_batch(std::string& err);
  bool run(const float* warm, long wld, int judge, std::string& err);
  bool fetch(double* X, double* Y, double* Sx, long* iters, long* status, long* fallback, double* resid,
             std::string& err);
  template <typename T>
  void fill_args(FamArgs<T>& a);
  template <typename T>
  FamArgs<T> args(int set_index, int ld, int cnt);
  template <typename T>
  bool run_t(const float* warm, long wld, int judge, std::string& err);
  template <typename T>
  bool iterate(FamArgs<T>& a, int steps, std::string& err);
  template <typename T>
  bool apply_A(FamArgs<T>& a, const T* Xin, T* Z);
  template <typename T>
  bool apply_At(FamArgs<T>& a, const T* X1, const T* X2, T* Z1, T* Z2);
  template <typename T>
  void prologue(FamArgs<T>& a);
  template <typename T>
  bool at_phase_b(FamArgs<T>& a);
  // chunks of phase B partials: one per A' task block (sparse, fused into the
  // A' product) or per RPC rows (dense)
  int b_chunks() const { return dense ? (n + RPC - 1) / RPC : hybrid() ? np / 64 : curAt->nAt; }
};

// Rows of a CSR matrix into SpMM wave tasks, NWV per block (k_fam_spmm):
// runs of complete rows of at most `te` entries (and 256 rows), or, for a
// row longer than `te`, one block whose NWV tasks split the row's entries.
static std::vector<int4> spmm_tasks(const std::vector<int>& ptr, int te) {
  std::vector<int4> out, cur;
  auto flush = [&]() {
    if (cur.empty()) return;
    while (cur.size() < (size_t)NWV) cur.push_back(int4{0, 0, 0, 0});
    out.insert(out.end(), cur.begin(), cur.end());
    cur.clear();
  };
  const int nrows = (int)ptr.size() - 1;
  int r = 0;
  while (r < nrows) {
    const int len = ptr[r + 1] - ptr[r];
    if (len > te) {
      flush();
      const int part = (len + NWV - 1) / NWV;
      for (int k = 0; k < NWV; ++k) {
        const int q0 = std::min(ptr[r + 1], ptr[r] + k * part), q1 = std::min(ptr[r + 1], q0 + part);
        out.push_back(int4{r, -1, q0, q1});
      }
      ++r;
      continue;
    }
    int e = r;
    while (e < nrows && e - r < 256 && ptr[e + 1] - ptr[r] <= te) ++e;
    cur.push_back(int4{r, e, ptr[r], ptr[e]});
    if (cur.size() == (size_t)NWV) flush();
    r = e;
  }
  flush();
  if (out.empty()) out.assign(NWV, int4{0, 0, 0, 0});
  return out;
}

static std::vector<int> entry_rows(const std::vector<int>& ptr) {
  std::vector<int> rows(ptr.empty() ? 0 : ptr.back());
  for (size_t r = 0; r + 1 < ptr.size(); ++r)
    for (int q = ptr[r]; q < ptr[r + 1]; ++q) rows[q] = (int)r;
  if (rows.empty()) rows.push_back(0);
  return rows;
}

bool Family::create(const Data* d, const Cone* k, const PdosOptions& o, std::string& err) {
  fp64 = o.precision == PDOS_FP64;
  if (o.bf16_products >= 1 && o.bf16_products <= 3) nprod = o.bf16_products;
  device = o.device;
  if (o.equil != PDOS_EQUIL_REFERENCE) {  // Ruiz scaling depends on each instance's b and c
    err = "the family engine shares one scaling of A across instances: PDOS_EQUIL_REFERENCE only";
    return false;
  }
  if (!build_setup(d, k, S, err)) return false;
  par = *d->p;
  n = S.n;
  m = S.m;
  np = pad64(n);
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      num_cu = prop.multiProcessorCount;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&fev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&fev[1], hipEventDisableTiming) != hipSuccess) {
    err = "device/stream setup failed";
    return false;
  }
  if (rocblas_create_handle(&blas) != rocblas_status_success ||
      rocblas_set_stream(blas, st) != rocblas_status_success) {
    err = "rocblas handle";
    return false;
  }
  rp = upload(S.rp, err); ri = upload(S.ri, err); rx = upload_vals(S.rx, err);
  cp = upload(S.cp, err); ci = upload(S.ci, err); cx = upload_vals(S.cx, err);
  soc = upload(S.soc_ptr, err); D = upload(S.D, err); E = upload(S.E, err); E0 = upload(S.E0, err);
  bshared = upload_vals(S.b, err);
  if (!err.empty()) return false;
  for (int gi = 0; gi < 2; ++gi) {
    Tasks& T = tasks[gi];
    const int te = gi == 0 ? 64 : 16;
    const std::vector<int4> tA = spmm_tasks(S.rp, te), tAt = spmm_tasks(S.cp, te);
    T.nA = (int)tA.size() / NWV;
    T.nAt = (int)tAt.size() / NWV;
    T.A = static_cast<int4*>(upload(tA, err));
    T.At = static_cast<int4*>(upload(tAt, err));
  }
  rowA = static_cast<int*>(upload(entry_rows(S.rp), err));
  rowAt = static_cast<int*>(upload(entry_rows(S.cp), err));
  std::function<int2(int)> rec_rows;  // set when A has dense rows
  std::vector<int> hc;  // dense rows of A (mixed precision, sparse A), ascending
  if (!fp64 && S.nnz() <= (long)m * n / 4) {
    for (int r = 0; r < m; ++r)
      if (S.rp[r + 1] - S.rp[r] >= kDenseRow) hc.push_back(r);
    std::stable_sort(hc.begin(), hc.end(),
                     [&](int x, int y) { return S.rp[x + 1] - S.rp[x] > S.rp[y + 1] - S.rp[y]; });
    if (hc.size() > (size_t)kDenseMax) hc.resize(kDenseMax);
    std::sort(hc.begin(), hc.end());
  }
  nh = (int)hc.size();
  if (nh > 0) {
    std::vector<int> sl(m, -1);
    for (int k = 0; k < nh; ++k) {
      sl[hc[k]] = k;
      drows.row[k] = hc[k];
    }
    // A'_h (np x kDenseMax, A'_h[i][k] = A[hc[k]][i]) in MFMA operand order
    std::vector<float> ah((size_t)np * kDenseMax, 0.f);
    for (int k = 0; k < nh; ++k)
      for (int q = S.rp[hc[k]]; q < S.rp[hc[k] + 1]; ++q) {
        const int i = S.ri[q];
        const int g = i / 16, il = i % 16, sq = k / 4, kq = k % 4;
        ah[((size_t)g * 8 + sq) * 64 + kq * 16 + il] += float(S.rx[q]);
      }
    std::vector<int> lp{0}, lidx;
    std::vector<double> lval;
    for (int i = 0; i < n; ++i) {
      for (int q = S.cp[i]; q < S.cp[i + 1]; ++q)
        if (sl[S.ci[q]] < 0) {
          lidx.push_back(S.ci[q]);
          lval.push_back(S.cx[q]);
        }
      lp.push_back((int)lidx.size());
    }
    const std::vector<int> lr = entry_rows(lp);
    std::vector<int4> lt;
    for (int t = 0; t < np / 64; ++t)
      for (int w = 0; w < NWV; ++w) {
        const int r0 = std::min(n, 64 * t + 8 * w), r1 = std::min(n, r0 + 8);
        lt.push_back(int4{r0, r1, lp[r0], lp[r1]});
      }
    if (lidx.empty()) {
      lidx.push_back(0);
      lval.push_back(0.0);
    }
    // A_h for the A product (32 x np, MFMA order) and the light CSR of A
    std::vector<float> ahr((size_t)np * kDenseMax, 0.f);
    for (int k = 0; k < nh; ++k)
      for (int q = S.rp[hc[k]]; q < S.rp[hc[k] + 1]; ++q) {
        const int j = S.ri[q];
        ahr[((size_t)(j / 4) * 2 + k / 16) * 64 + (j % 4) * 16 + k % 16] += float(S.rx[q]);
      }
    std::vector<int> lrp{0}, lridx;
    std::vector<double> lrval;
    for (int r = 0; r < m; ++r) {
      if (sl[r] < 0)
        for (int q = S.rp[r]; q < S.rp[r + 1]; ++q) {
          lridx.push_back(S.ri[q]);
          lrval.push_back(S.rx[q]);
        }
      lrp.push_back((int)lridx.size());
    }
    rec_rows = [&, sl, lrp, lridx, lrval](int j) {  // phase C's record of row j of A
      if (sl[j] >= 0) return int2{-1, sl[j]};
      const int q0 = lrp[j], q1 = lrp[j + 1];
      if (q1 - q0 != 1) return int2{q1 == q0 ? -2 : -3, 0};
      const float v = float(lrval[q0]);
      int vb;
      std::memcpy(&vb, &v, 4);
      return int2{lridx[q0], vb};
    };
    if (lridx.empty()) {
      lridx.push_back(0);
      lrval.push_back(0.0);
    }
    std::vector<float> lrvf(lrval.begin(), lrval.end());
    Ahr = static_cast<float*>(upload(ahr, err));
    lrpd = static_cast<int*>(upload(lrp, err));
    lri = static_cast<int*>(upload(lridx, err));
    lrxf = static_cast<float*>(upload(lrvf, err));
    Ah = static_cast<float*>(upload(ah, err));
    li = static_cast<int*>(upload(lidx, err));
    lx = upload_vals(lval, err);
    lrow = static_cast<int*>(upload(lr, err));
    ltasks = static_cast<int4*>(upload(lt, err));
  }
  {
    std::vector<int4> v1, v2;
    std::vector<int> vc;
    for (int i = 0; i < n; i += RPC) v1.push_back(int4{i, std::min(n, i + RPC), 0, 0});
    nx_items = (int)v1.size();
    for (int j = 0; j < S.f + S.l; j += RPC) v1.push_back(int4{j, std::min(S.f + S.l, j + RPC), 1, 0});
    nlin_items = (int)v1.size() - nx_items;
    for (int cn = 0; cn < S.nsoc(); ++cn) {
      const int r0 = S.soc_ptr[cn], r1 = S.soc_ptr[cn + 1];
      const int g0 = (int)v1.size();
      for (int j = r0; j < r1; j += RPC) v1.push_back(int4{j, std::min(r1, j + RPC), 2, cn});
      const int ng = (int)v1.size() - g0;
      for (int g = g0; g < g0 + ng; ++g) {
        v2.push_back(int4{v1[g].x, v1[g].y, g0, ng});
        vc.push_back(cn);
      }
    }
    nc1 = (int)v1.size();
    nc2 = (int)v2.size();
    c1 = static_cast<int4*>(upload(v1, err));
    if (rec_rows) {
      // phase C's record for row j = item.x + w + NWV k of every item (RW per wave)
      std::vector<int2> r8((size_t)nc1 * NWV * RW, int2{-2, 0});
      for (int t = 0; t < nc1; ++t)
        if (v1[t].z != 0)
          for (int w = 0; w < NWV; ++w)
            for (int k = 0; k < RW; ++k) {
              const int j = v1[t].x + w + k * NWV;
              if (j < v1[t].y) r8[((size_t)t * NWV + w) * RW + k] = rec_rows(j);
            }
      rec8d = static_cast<int2*>(upload(r8, err));
    }
    if (!v2.empty()) {
      c2 = static_cast<int4*>(upload(v2, err));
      c2_cone = static_cast<int*>(upload(vc, err));
    }
  }
  if (!err.empty()) return false;
  // M = I + A'A of the scaled A on the device (rocBLAS syrk), then
  // G = M^-1 by Cholesky (rocSOLVER potrf + potri), once per family.
  std::vector<double> Acol((size_t)m * n, 0.0);  // column-major m x n
  for (int j = 0; j < n; ++j)
    for (int p = S.cp[j]; p < S.cp[j + 1]; ++p) Acol[(size_t)j * m + S.ci[p]] += S.cx[p];
  std::vector<double> M((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) M[(size_t)i * n + i] = 1.0;
  double* dA = static_cast<double*>(dalloc(Acol.size() * 8, allocs, err));
  double* dM = static_cast<double*>(dalloc(M.size() * 8, allocs, err));
  rocblas_int* dinfo = static_cast<rocblas_int*>(dalloc(sizeof(rocblas_int), allocs, err));
  if (!dA || !dM || !dinfo) return false;
  (void)hipMemcpyAsync(dA, Acol.data(), Acol.size() * 8, hipMemcpyHostToDevice, st);
  (void)hipMemcpyAsync(dM, M.data(), M.size() * 8, hipMemcpyHostToDevice, st);
  const double one = 1.0;
  if (rocblas_dsyrk(blas, rocblas_fill_lower, rocblas_operation_transpose, n, m, &one, dA, m, &one, dM, n) !=
          rocblas_status_success ||
      rocsolver_dpotrf(blas, rocblas_fill_lower, n, dM, n, dinfo) != rocblas_status_success ||
      rocsolver_dpotri(blas, rocblas_fill_lower, n, dM, n, dinfo) != rocblas_status_success) {
    err = "rocblas syrk / rocsolver potrf, potri failed";
    return false;
  }
  // dense A: the products run as GEMMs on a row-major copy (column-major n x m)
  dense = S.nnz() > (long)m * n / 4;
  if (dense) {
    std::vector<double> Arow((size_t)m * n, 0.0);
    for (int r = 0; r < m; ++r)
      for (int p = S.rp[r]; p < S.rp[r + 1]; ++p) Arow[(size_t)r * n + S.ri[p]] += S.rx[p];
    Ad = upload_vals(Arow, err);
    if (!Ad) return false;
    split_dense = !fp64;
    if (split_dense) {
      m64 = (m + 63) & ~63;
      n64 = (n + 63) & ~63;
      // rows padded to the products' tile height, columns to 64
      std::vector<__bf16> th((size_t)dense_row_pad(n) * m64, (__bf16)0.f), tl(th),
          rh((size_t)dense_row_pad(m) * n64, (__bf16)0.f), rl(rh);
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < n; ++c) {
          const float v = float(Arow[(size_t)r * n + c]);
          const __bf16 hi = (__bf16)v, lo = (__bf16)(v - float(hi));
          th[(size_t)c * m64 + r] = hi;
          tl[(size_t)c * m64 + r] = lo;
          rh[(size_t)r * n64 + c] = hi;
          rl[(size_t)r * n64 + c] = lo;
        }
      ATh = upload(th, err);
      ATl = upload(tl, err);
      ARh = upload(rh, err);
      ARl = upload(rl, err);
      if (!ATh || !ATl || !ARh || !ARl) return false;
    }
  }
  rocblas_int info = 0;
  (void)hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st);
  (void)hipStreamSynchronize(st);
  if (info != 0) { err = "I + A'A is not positive definite"; return false; }
  if (fp64) {
    Gd = dalloc((size_t)n * n * 8, allocs, err);
    if (!Gd) return false;
    hipLaunchKernelGGL(k_symmetrize_convert<double>, dim3(1024), dim3(256), 0, st, n, dM,
                       static_cast<double*>(Gd), n);
  } else {
    Gcat = dalloc((size_t)np * 3 * np * 2, allocs, err);
    if (!Gcat) return false;
    hipLaunchKernelGGL(k_split_g, dim3(1024), dim3(256), 0, st, n, np, dM, static_cast<__bf16*>(Gcat));
  }
  return hipStreamSynchronize(st) == hipSuccess;
}

// Per-instance c (count x n) and optional b (count x m), original units:
// scaled exactly as build_setup does (common.h:140-192) and uploaded
// feature-major together with lambda, pscale, dscale.
bool Family::load(int cnt, const double* C, const double* B, std::string& err) {
  // a batch of the same padded width (and kind of b) keeps its device
  // buffers: every run initializes its working sets from the loaded inputs
  const bool reuse = hflags && count > 0 && pad64(cnt) == ldb && b_inst == (B != nullptr);
  if (!reuse) free_batch();
  count = cnt;
  ldb = pad64(cnt);
  b_inst = B != nullptr;
  const size_t tsz = fp64 ? 8 : 4;
  if (!reuse && !alloc_batch(err)) return false;
  // the raw inputs to the device; setup per instance there (k_fam_scale_in)
  (void)hipMemcpyAsync(inC, C, (size_t)cnt * n * 8, hipMemcpyHostToDevice, st);
  if (B) (void)hipMemcpyAsync(inB, B, (size_t)cnt * m * 8, hipMemcpyHostToDevice, st);
  // shared b: its stop-test scale 1 + |b| (pdos.c:76) and |b~|^2 of the
  // scaled b, summed in index order as per instance
  double nbs_sh = 0.0;
  for (int j = 0; j < m; ++j) nbs_sh += S.b[j] * S.b[j];
  // one wave per block: a lane walks its instance's row of C, so blocks of
  // 64 spread the batch over 4x as many CUs as blocks of 256 would
  const dim3 grid((ldb + 63) / 64);
  if (fp64)
    hipLaunchKernelGGL(k_fam_scale_in<double>, grid, dim3(64), 0, st, cnt, n, m, ldb, static_cast<const double*>(inC),
                       static_cast<const double*>(inB), static_cast<const double*>(E0),
                       static_cast<const double*>(D), S.pscale, nbs_sh, static_cast<double*>(base_c),
                       static_cast<double*>(base_b), static_cast<double*>(base_lam), static_cast<double*>(base_ps),
                       static_cast<double*>(base_ds));
  else
    hipLaunchKernelGGL(k_fam_scale_in<float>, grid, dim3(64), 0, st, cnt, n, m, ldb, static_cast<const double*>(inC),
                       static_cast<const double*>(inB), static_cast<const double*>(E0),
                       static_cast<const double*>(D), S.pscale, nbs_sh, static_cast<float*>(base_c),
                       static_cast<float*>(base_b), static_cast<double*>(base_lam), static_cast<double*>(base_ps),
                       static_cast<double*>(base_ds));
  if (hipGetLastError() != hipSuccess) { err = "load: scaling kernel failed"; return false; }
  (void)tsz;
  return hipStreamSynchronize(st) == hipSuccess;
}

// Device buffers of a batch of width ldb (count and b_inst set).
bool Family::alloc_batch(std::string& err) {
  const size_t tsz = fp64 ? 8 : 4;
  inC = dalloc((size_t)ldb * n * 8, batch_allocs, err);
  inB = b_inst ? dalloc((size_t)ldb * m * 8, batch_allocs, err) : nullptr;
  gout = dalloc((size_t)ldb * (n + 2 * m) * 8, batch_allocs, err);
  base_c = dalloc((size_t)n * ldb * tsz, batch_allocs, err);
  base_b = b_inst ? dalloc((size_t)m * ldb * tsz, batch_allocs, err) : nullptr;
  base_lam = dalloc((size_t)ldb * 8, batch_allocs, err);
  base_ps = dalloc((size_t)ldb * 8, batch_allocs, err);
  base_ds = dalloc((size_t)ldb * 8, batch_allocs, err);
  auto vec = [&](int rows) { return dalloc((size_t)rows * ldb * tsz, batch_allocs, err); };
  for (int k = 0; k < 2; ++k) {
    Set& q = set[k];
    q.X = vec(np); q.XN = vec(np); q.S = vec(m); q.AX = vec(m);
    q.Y = vec(2 * m);
    q.W = q.Y ? static_cast<unsigned char*>(q.Y) + (size_t)m * ldb * tsz : nullptr;
    q.U = vec(2 * n);
    q.V = q.U ? static_cast<unsigned char*>(q.U) + (size_t)n * ldb * tsz : nullptr;
    q.R = vec(np);
    q.Rcat = fp64 ? nullptr : dalloc((size_t)3 * np * ldb * 2, batch_allocs, err);
    q.c = vec(n);
    q.b = b_inst ? vec(m) : nullptr;
    q.lam = dalloc((size_t)ldb * 8, batch_allocs, err);
    q.ps = dalloc((size_t)ldb * 8, batch_allocs, err);
    q.ds = dalloc((size_t)ldb * 8, batch_allocs, err);
    q.ints = dalloc((size_t)4 * ldb * 4, batch_allocs, err);
    q.red = dalloc((size_t)8 * ldb * 8, batch_allocs, err);
    q.orig = dalloc((size_t)ldb * 4, batch_allocs, err);
    q.pa = dalloc((size_t)std::max(1, m_items()) * ldb * 2 * 8, batch_allocs, err);
    const int nb = std::max({(n + RPC - 1) / RPC, tasks[0].nAt, tasks[1].nAt});
    q.pb = dalloc((size_t)nb * ldb * 2 * 8, batch_allocs, err);
    q.pc = dalloc((size_t)nc1 * ldb * 8, batch_allocs, err);
    q.pt0 = dalloc((size_t)std::max(1, S.nsoc()) * ldb * 8, batch_allocs, err);
  }
  Xo = vec(n); Yo = vec(m); So = vec(m);
  if (split_dense) {  // planes for the widest split of any width the batch compacts to
    size_t cap = 0;
    for (int w = 64; w <= ldb; w += 64) cap = std::max(cap, (size_t)dense_at_splits(n, m64, w) * w * n);
    Upl = dalloc(cap * 4, batch_allocs, err);
    Vpl = dalloc(cap * 4, batch_allocs, err);
  }
  Io = dalloc((size_t)3 * ldb * 4, batch_allocs, err);
  Ro = dalloc((size_t)8 * ldb * 8, batch_allocs, err);
  dlist = static_cast<int*>(dalloc((size_t)ldb * 4, batch_allocs, err));
  AXp = hybrid() ? static_cast<float*>(dalloc((size_t)(np / 64) * 32 * ldb * 4, batch_allocs, err)) : nullptr;
  if (hipHostMalloc(reinterpret_cast<void**>(&hflags), (size_t)2 * ldb * 4, hipHostMallocDefault) != hipSuccess) {
    hflags = nullptr;
    err = "hipHostMalloc failed";
  }
  if (!err.empty()) return false;
  return hipStreamSynchronize(st) == hipSuccess;
}

template <typename T>
FamArgs<T> Family::args(int k, int ld, int cnt) {
  FamArgs<T> a;
  fill_args(a);
  const Set& q = set[k];
  a.ldb = ld;
  a.count = cnt;
  a.b = b_inst ? static_cast<T*>(q.b) : static_cast<T*>(bshared);
  a.c = static_cast<T*>(q.c);
  a.lam = static_cast<double*>(q.lam);
  a.pscale = static_cast<double*>(q.ps);
  a.dscale = static_cast<double*>(q.ds);
  a.X = static_cast<T*>(q.X); a.XN = static_cast<T*>(q.XN); a.S = static_cast<T*>(q.S);
  a.Y = static_cast<T*>(q.Y); a.AX = static_cast<T*>(q.AX); a.W = static_cast<T*>(q.W);
  a.U = static_cast<T*>(q.U); a.V = static_cast<T*>(q.V);
  a.R = static_cast<T*>(q.R);
  a.Rcat = static_cast<__bf16*>(q.Rcat);
  int* ip = static_cast<int*>(q.ints);
  a.done = ip; a.status = ip + ld; a.iters = ip + 2 * ld; a.fallback = ip + 3 * ld;
  a.red = static_cast<double*>(q.red);
  a.orig = static_cast<int*>(q.orig);
  a.pa = static_cast<double*>(q.pa);
  a.pb = static_cast<double*>(q.pb);
  a.pc = static_cast<double*>(q.pc);
  a.pt0 = static_cast<double*>(q.pt0);
  return a;
}

template <typename T>
void Family::fill_args(FamArgs<T>& a) {
  std::memset(&a, 0, sizeof(a));
  a.n = n; a.m = m; a.f = S.f; a.lend = S.f + S.l; a.nsoc = S.nsoc();
  a.ldb = ldb; a.count = count; a.np = np; a.b_inst = b_inst;
  a.max_iters = (int)std::min<long>(std::max<long>(par.MAX_ITERS, 0), 0x7fffffff);
  a.eps = par.EPS_ABS; a.relax = par.ALPHA;
  a.soc_ptr = static_cast<const int*>(soc);
  a.rp = static_cast<const int*>(rp); a.ri = static_cast<const int*>(ri); a.rx = static_cast<const T*>(rx);
  a.cp = static_cast<const int*>(cp); a.ci = static_cast<const int*>(ci); a.cx = static_cast<const T*>(cx);
  a.D = static_cast<const double*>(D); a.E = static_cast<const double*>(E);
  a.nx_items = nx_items;
  a.nlin_items = nlin_items;
  a.nprod = nprod;
  a.nsplit = 1;
  a.plane = 0;
  if (hybrid()) {
    a.rec8 = rec8d;
    a.AXp = AXp;
    a.mt = np / 64;
    a.lrp = lrpd;
    a.lri = lri;
    a.lrx = lrxf;
  }
}

template <typename T>
bool Family::iterate(FamArgs<T>& a, int steps, std::string& err) {
  const int cols = a.ldb / LANES;
  // SpMM tasks: 64 entries while the batch is large, 16 once it is small
  curA = &tasks[pick_tasks(cols, tasks[0].nA, tasks[1].nA)];
  curAt = &tasks[pick_tasks(cols, tasks[0].nAt, tasks[1].nAt)];
  const int cm = m_items(), cn = b_chunks();
  const dim3 block(FT);
  cur_ld = a.ldb;
  for (int s = 0; s < steps; ++s) {
    mark(K_AT);
    if (!at_phase_b(a)) { err = "A' product failed"; return false; }
    if (dense) {
      mark(K_B);
      hipLaunchKernelGGL(k_fam_decide<T>, dim3(cols), block, 0, st, a, cm, cn);
    }
    mark(K_GEMM);
    if (fp64) {
      const double one = 1.0, zero = 0.0;
      // row-major XN (n x ldb) = G (n x n) R (n x ldb): column-major XN' = R' G
      if (rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, a.ldb, n, n, &one,
                        reinterpret_cast<const double*>(a.R), a.ldb, static_cast<const double*>(Gd), n, &zero,
                        reinterpret_cast<double*>(a.XN), a.ldb) != rocblas_status_success) {
        err = "rocblas_dgemm failed";
        return false;
      }
    } else {
      // A with dense rows: the launch also forms their AX planes and runs
      // the stop test; phase C forms the rest of AX
      const int mt = np / 64, nt = a.ldb / 64;
      const bool fuse = hybrid();
      FamArgs<float> fa{};
      if constexpr (std::is_same<T, float>::value) fa = a;
      hipLaunchKernelGGL(k_fam_xupdate, dim3(mt * nt + (fuse ? nt : 0)), dim3(256), 0, st, mt, nt, nprod * np / 32,
                         static_cast<const __bf16*>(Gcat), 3 * np, a.Rcat, a.ldb,
                         reinterpret_cast<const float*>(a.X), reinterpret_cast<float*>(a.XN),
                         fuse ? Ahr : nullptr, fuse ? AXp : nullptr, fa, cm, cn);
    }
    if (dense) {
      mark(K_AX);
      if (!apply_A(a, a.XN, a.AX)) { err = "A product failed"; return false; }
    } else if (hybrid()) {
      // formed by the x-update and phase C
    } else {
      mark(K_AX);
      const SpmmIn<T, 1> in{curA->A, rowA, a.ri, a.rx, a.XN, a.XN, a.XN, a.ldb, cols, curA->nA};
      hipLaunchKernelGGL(k_fam_ax_decide<T>, dim3(cols + cols * curA->nA), block, 0, st, a, in, cm, cn);
    }
    mark(K_C);
    bool ll1 = false;
    if constexpr (std::is_same<T, float>::value) {
      if (hybrid() && cols <= kNarrowCols) {
        hipLaunchKernelGGL(k_fam_c1_ll<float>, dim3(cols, nc1), block, 0, st, a, c1);
        ll1 = true;
      }
    }
    if (!ll1) hipLaunchKernelGGL(k_fam_c1<T>, dim3(cols, nc1), block, 0, st, a, c1);
    if (nc2) {
      bool ll = false;
      if constexpr (std::is_same<T, float>::value) {
        if (hybrid() && cols <= kNarrowCols) {
          hipLaunchKernelGGL(k_fam_c2_ll<float>, dim3(cols, nc2), block, 0, st, a, c2, c2_cone);
          ll = true;
        }
      }
      if (!ll) hipLaunchKernelGGL(k_fam_c2<T>, dim3(cols, nc2), block, 0, st, a, c2, c2_cone);
    }
    std::swap(a.X, a.XN);
    mark(-1);
  }
  return hipGetLastError() == hipSuccess;
}

// Z = A X (m x ldb) for the batch: sparse rows or a dense GEMM.
template <typename T>
bool Family::apply_A(FamArgs<T>& a, const T* Xin, T* Z) {
  if (!dense) {
    const int cols = a.ldb / LANES;
    const SpmmIn<T, 1> in{curA->A, rowA, a.ri, a.rx, Xin, Xin, Xin, a.ldb, cols, curA->nA};
    hipLaunchKernelGGL((k_fam_spmm<T, 1>), dim3(cols * curA->nA), dim3(FT), 0, st, in, Z, Z, Z);
    return hipGetLastError() == hipSuccess;
  }
  const T one = T(1), zero = T(0);
  if constexpr (std::is_same<T, float>::value) {
    if (split_dense)  // Z = A X, split bf16 MFMA (dense_gemm.hip)
      return dense_a(st, m, n, a.ldb, n64, static_cast<const __bf16*>(ARh), static_cast<const __bf16*>(ARl), Xin, Z) ==
             hipSuccess;
  }
  // column-major Z' (ldb x m) = X' (ldb x n) * A^T, A^T stored column-major n x m (ld n)
  if constexpr (std::is_same<T, double>::value)
    return rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, a.ldb, m, n, &one, Xin, a.ldb,
                         static_cast<const double*>(Ad), n, &zero, Z, a.ldb) == rocblas_status_success;
  else
    return rocblas_sgemm(blas, rocblas_operation_none, rocblas_operation_none, a.ldb, m, n, &one, Xin, a.ldb,
                         static_cast<const float*>(Ad), n, &zero, Z, a.ldb) == rocblas_status_success;
}

// Z1 = A' X1 and Z2 = A' X2 (n x ldb): one sparse pass or one dense GEMM
// per right-hand side.
template <typename T>
bool Family::apply_At(FamArgs<T>& a, const T* X1, const T* X2, T* Z1, T* Z2) {
  if (!dense) {
    const int cols = a.ldb / LANES;
    const SpmmIn<T, 2> in{curAt->At, rowAt, a.ci, a.cx, X1, X2, X2, a.ldb, cols, curAt->nAt};
    hipLaunchKernelGGL((k_fam_spmm<T, 2>), dim3(cols * curAt->nAt), dim3(FT), 0, st, in, Z1, Z2, Z2);
    return hipGetLastError() == hipSuccess;
  }
  const T one = T(1), zero = T(0);
  // column-major Z_k' (ldb x n) = X_k' (ldb x m) * A, A = (stored n x m)^T:
  // NT GEMMs (measured 3x faster here than NN on a column-major copy of A),
  // both right-hand sides in one strided-batched launch (X2 and Z2 lie at
  // fixed offsets after X1 and Z1 in one allocation each), which gives the
  // launch twice the output tiles: a 1,024-column A' X' has ~128 of them
  // per product, fewer than the GPU's 256 CUs
  const rocblas_stride sx = X2 - X1, sz = Z2 - Z1;
  rocblas_status r;
  if constexpr (std::is_same<T, double>::value)
    r = rocblas_dgemm_strided_batched(blas, rocblas_operation_none, rocblas_operation_transpose, a.ldb, n, m, &one,
                                      X1, a.ldb, sx, static_cast<const double*>(Ad), n, 0, &zero, Z1, a.ldb, sz, 2);
  else
    r = rocblas_sgemm_strided_batched(blas, rocblas_operation_none, rocblas_operation_transpose, a.ldb, n, m, &one,
                                      X1, a.ldb, sx, static_cast<const float*>(Ad), n, 0, &zero, Z1, a.ldb, sz, 2);
  return r == rocblas_status_success;
}

// A' products and phase B: for a sparse A one fused pass (k_fam_at_b), for a
// dense A the GEMMs then k_fam_b.
template <typename T>
bool Family::at_phase_b(FamArgs<T>& a) {
  if (!dense) {
    const int cols = a.ldb / LANES;
    if constexpr (std::is_same<T, float>::value) {
      if (hybrid()) {
        const SpmmIn<T, 2> in{ltasks, lrow, li, static_cast<const T*>(lx), a.Y, a.W, a.W, a.ldb, cols, np / 64};
        hipLaunchKernelGGL(k_fam_at_hyb<T>, dim3(cols * (np / 64)), dim3(FT), 0, st, a, in, Ah, drows);
        return hipGetLastError() == hipSuccess;
      }
    }
    const SpmmIn<T, 2> in{curAt->At, rowAt, a.ci, a.cx, a.Y, a.W, a.W, a.ldb, cols, curAt->nAt};
    hipLaunchKernelGGL(k_fam_at_b<T>, dim3(cols * curAt->nAt), dim3(FT), 0, st, a, in);
    return hipGetLastError() == hipSuccess;
  }
  bool done_at = false;
  if constexpr (std::is_same<T, float>::value) {
    if (split_dense) {  // U, V = A'Y, A'W as split-K planes (dense_gemm.hip), summed by phase B
      const int sp = dense_at_splits(n, m64, a.ldb);
      a.U = static_cast<float*>(Upl);
      a.V = static_cast<float*>(Vpl);
      a.nsplit = sp;
      a.plane = (long)n * a.ldb;
      if (dense_at_pair(st, m, n, a.ldb, m64, static_cast<const __bf16*>(ATh), static_cast<const __bf16*>(ATl), a.Y,
                        a.W, a.U, a.V, (size_t)a.plane, sp) != hipSuccess)
        return false;
      done_at = true;
    }
  }
  if (!done_at && !apply_At(a, a.Y, a.W, a.U, a.V)) return false;
  const dim3 gn(a.ldb / LANES, (n + RPC - 1) / RPC);
  if (fp64) hipLaunchKernelGGL((k_fam_b<T, false>), gn, dim3(FT), 0, st, a);
  else hipLaunchKernelGGL((k_fam_b<T, true>), gn, dim3(FT), 0, st, a);
  return hipGetLastError() == hipSuccess;
}

// W and the stop-test partials of the current iterate (phase C forms them
// inside the loop).
template <typename T>
void Family::prologue(FamArgs<T>& a) {
  cur_ld = a.ldb;
  mark(K_A);
  hipLaunchKernelGGL(k_fam_a<T>, dim3(a.ldb / LANES, m_items()), dim3(FT), 0, st, a, c1 + nx_items);
  mark(-1);
  need_prologue = false;
}

template <typename T>
bool Family::run_t(const float* warm, long wld, int judge, std::string& err) {
  const size_t tsz = sizeof(T);
  int cur = 0, ld = ldb, cnt = count;
  FamArgs<T> a = args<T>(0, ld, cnt);
  // constants of the loaded batch into working set 0, identity column map
  (void)hipMemcpyAsync(a.c, base_c, (size_t)n * ld * tsz, hipMemcpyDeviceToDevice, st);
  if (b_inst) (void)hipMemcpyAsync(a.b, base_b, (size_t)m * ld * tsz, hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(a.lam, base_lam, (size_t)ld * 8, hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(a.pscale, base_ps, (size_t)ld * 8, hipMemcpyDeviceToDevice, st);
  (void)hipMemcpyAsync(a.dscale, base_ds, (size_t)ld * 8, hipMemcpyDeviceToDevice, st);
  hipLaunchKernelGGL(k_iota, dim3(64), dim3(256), 0, st, ld, a.orig);
  const dim3 block(FT);
  curA = &tasks[pick_tasks(ld / LANES, tasks[0].nA, tasks[1].nA)];
  curAt = &tasks[pick_tasks(ld / LANES, tasks[0].nAt, tasks[1].nAt)];
  hipLaunchKernelGGL(k_fam_init<T>, dim3(ld / LANES), block, 0, st, a, warm, wld);
  if (!apply_A(a, a.X, a.AX)) { err = "A product failed"; return false; }
  need_prologue = true;
  if (warm && judge) {
    const int cm = m_items(), cn = b_chunks();
    a.judge = 1;
    prologue(a);
    if (!at_phase_b(a)) { err = "A' product failed"; return false; }
    hipLaunchKernelGGL(k_fam_decide<T>, dim3(ld / LANES), block, 0, st, a, cm, cn);
    a.judge = 0;
    need_prologue = true;  // rejected instances were reset
  }
  // Chunks of 16 iterations; after each, its done flags are copied to pinned
  // host memory.  Outside timing runs the next chunk is launched before the
  // host reads the flags (one chunk ahead: the GPU does not wait for the
  // check); the run-ahead chunk only repeats finished columns' frozen state.
  std::vector<int> list;
  compactions = 0;
  long total = 0;
  const bool ahead = !timing;
  // launches one chunk into flag slot `slot`: iterations launched, 0 when the
  // iteration budget is spent, -1 on error
  auto launch = [&](int slot) -> int {
    const int steps = (int)std::min<long>(16, a.max_iters + 2 - total);
    if (steps <= 0) return 0;
    if (need_prologue) prologue(a);
    if (!iterate(a, steps, err)) return -1;
    total += steps;
    (void)hipMemcpyAsync(hflags + (size_t)slot * ldb, a.done, (size_t)cnt * 4, hipMemcpyDeviceToHost, st);
    (void)hipEventRecord(fev[slot], st);
    return steps;
  };
  auto live = [&](const int* hd) {
    list.clear();
    for (int t = 0; t < cnt; ++t)
      if (!hd[t]) list.push_back(t);
    return (int)list.size();
  };
  int slot = 0;
  int r = launch(slot);
  while (r > 0) {
    const int nxt = ahead ? launch(slot ^ 1) : 0;
    if (nxt < 0) return false;
    if (hipEventSynchronize(fev[slot]) != hipSuccess) { err = "family engine failed"; return false; }
    collect();
    int active = live(hflags + (size_t)slot * ldb);
    if (active == 0) break;
    // compact once it drops at least 1/16 of the width: a move costs a few
    // microseconds, a dead column costs its share of every later iteration
    if (pad64(active) <= ld - std::max(LANES, ld / 16)) {
      if (nxt > 0) {  // the run-ahead chunk has moved on: compact its state
        if (hipStreamSynchronize(st) != hipSuccess) { err = "family engine failed"; return false; }
        active = live(hflags + (size_t)(slot ^ 1) * ldb);
        if (active == 0) break;
      }
      // compaction: finished columns to the output store, live ones to the other set
      hipLaunchKernelGGL(k_fam_scatter<T>, dim3(ld / LANES, kMoveSplit), block, 0, st, a, ldb, static_cast<T*>(Xo),
                         static_cast<T*>(Yo), static_cast<T*>(So), static_cast<int*>(Io), static_cast<double*>(Ro));
      (void)hipMemcpyAsync(dlist, list.data(), (size_t)active * 4, hipMemcpyHostToDevice, st);
      const int nld = pad64(active);
      FamArgs<T> b = args<T>(cur ^ 1, nld, active);
      hipLaunchKernelGGL(k_fam_move<T>, dim3(nld / LANES, kMoveSplit), block, 0, st, a, b, dlist, hybrid() ? m_items() : 0);
      (void)hipStreamSynchronize(st);  // dlist and the flag slots are reused
      cur ^= 1;
      ld = nld;
      cnt = active;
      a = b;
      ++compactions;
      need_prologue = !hybrid();  // with dense rows W and the partials were moved
      slot = 0;
      r = launch(slot);
      if (r < 0) return false;
      continue;
    }
    if (nxt > 0) {
      slot ^= 1;
      r = nxt;
    } else {
      r = launch(slot);
      if (r < 0) return false;
    }
  }
  if (r < 0) return false;
  hipLaunchKernelGGL(k_fam_scatter<T>, dim3(ld / LANES, kMoveSplit), block, 0, st, a, ldb, static_cast<T*>(Xo),
                     static_cast<T*>(Yo), static_cast<T*>(So), static_cast<int*>(Io), static_cast<double*>(Ro));
  return hipGetLastError() == hipSuccess;
}

bool Family::run(const float* warm, long wld, int judge, std::string& err) {
  if (hipSetDevice(device) != hipSuccess) {  // the calling thread's device
    err = "hipSetDevice failed";
    return false;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, st);
  const bool ok = fp64 ? run_t<double>(warm, wld, judge, err) : run_t<float>(warm, wld, judge, err);
  (void)hipEventRecord(e1, st);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  last_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ok;
}

bool Family::fetch(double* X, double* Y, double* Sx, long* iters, long* status, long* fallback,
                   double* resid, std::string& err) {
  const int cnt = count;
  std::vector<int> ip((size_t)3 * ldb);
  std::vector<double> red((size_t)8 * ldb);
  // unscaled on the device (x = E x, y = D y, s = s / D: pdos.c:251-294) into
  // instance-major binary64, then into the caller's arrays
  double* g = static_cast<double*>(gout);
  struct Part {
    const char* dev;
    char* out;
    size_t bytes;
  };
  std::vector<Part> parts;
  auto gather = [&](const void* Zp, int rows, const void* scale, int divide, double* dev, double* out) {
    if (!out) return;
    if (fp64)
      hipLaunchKernelGGL(k_gather_out<double>, dim3(1024), dim3(256), 0, st, rows, cnt, ldb,
                         static_cast<const double*>(Zp), static_cast<const double*>(scale), divide, dev);
    else
      hipLaunchKernelGGL(k_gather_out<float>, dim3(1024), dim3(256), 0, st, rows, cnt, ldb,
                         static_cast<const float*>(Zp), static_cast<const double*>(scale), divide, dev);
    parts.push_back(Part{reinterpret_cast<const char*>(dev), reinterpret_cast<char*>(out), (size_t)rows * cnt * 8});
  };
  gather(Xo, n, E, 0, g, X);
  gather(Yo, m, D, 0, g + (size_t)n * ldb, Y);
  gather(So, m, D, 1, g + (size_t)(n + m) * ldb, Sx);
  (void)hipMemcpyAsync(ip.data(), Io, (size_t)3 * ldb * 4, hipMemcpyDeviceToHost, st);
  (void)hipMemcpyAsync(red.data(), Ro, (size_t)8 * ldb * 8, hipMemcpyDeviceToHost, st);
  // the outputs in chunks of at most kFetchChunk bytes, through two pinned
  // stages: chunk c + 2 goes down while the host copies chunk c + 1 out
  // (pageable destinations take device copies at a fraction of the link's
  // rate; the host's copy is spread over up to 16 threads)
  if (!stage) {
    if (hipHostMalloc(reinterpret_cast<void**>(&stage), 2 * kFetchChunk, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&sev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sev[1], hipEventDisableTiming) != hipSuccess) {
      stage = nullptr;
      err = "fetch: pinned staging failed";
      return false;
    }
  }
  std::vector<Part> chunks;
  for (const Part& p : parts)
    for (size_t o = 0; o < p.bytes; o += kFetchChunk)
      chunks.push_back(Part{p.dev + o, p.out + o, std::min(kFetchChunk, p.bytes - o)});
  auto down = [&](size_t c) {
    char* buf = stage + (c & 1) * kFetchChunk;
    (void)hipMemcpyAsync(buf, chunks[c].dev, chunks[c].bytes, hipMemcpyDeviceToHost, st);
    (void)hipEventRecord(sev[c & 1], st);
  };
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (size_t c = 0; c < chunks.size() && c < 2; ++c) down(c);
  for (size_t c = 0; c < chunks.size(); ++c) {
    if (hipEventSynchronize(sev[c & 1]) != hipSuccess) {
      err = "fetch failed";
      return false;
    }
    const char* src = stage + (c & 1) * kFetchChunk;
    const size_t bytes = chunks[c].bytes, share = (bytes / nth + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) {
      const size_t o = (size_t)t * share;
      if (o >= bytes) break;
      char* dst = chunks[c].out + o;
      const char* from = src + o;
      const size_t len = std::min(share, bytes - o);
      th.emplace_back([dst, from, len] { std::memcpy(dst, from, len); });
    }
    std::memcpy(chunks[c].out, src, std::min(share, bytes));
    for (std::thread& t : th) t.join();
    if (c + 2 < chunks.size()) down(c + 2);
  }
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) {
    err = "fetch failed";
    return false;
  }
  for (int t = 0; t < cnt; ++t) {
    if (iters) iters[t] = ip[t];
    if (status) status[t] = ip[ldb + t];
    if (fallback) fallback[t] = ip[2 * ldb + t];
    if (resid)
      for (int q = 0; q < 5; ++q) resid[(size_t)t * 5 + q] = red[(size_t)(2 + q) * ldb + t];
  }
  return true;
}

}  // namespace pdos_amd

// A family handle: one engine, or for mixed precision two, which run the two
// halves of a large batch at once on their own streams (host threads): the
// narrow, latency-bound end of one half overlaps the other's.
struct PdosFamily {
  std::vector<std::unique_ptr<pdos_amd::Family>> F;
  std::vector<int> first, count;  // instance range of each engine in use
  int max_engines = 2;            // engines a batch is split over (pdos_family_engines)
};

namespace {
thread_local std::string g_fam_err;
constexpr int kSplitMin = 8192;  // smallest batch split over two engines
}

extern "C" {

const char* pdos_family_last_error(void) { return g_fam_err.c_str(); }

PdosFamily* pdos_family_create(const Data* d, const Cone* k, const PdosOptions* opt) {
  PdosOptions o;
  if (opt) o = *opt;
  else pdos_default_options(&o);
  PdosFamily* h = new PdosFamily;
  const int engines = o.precision == PDOS_FP64 ? 1 : 2;
  for (int e = 0; e < engines; ++e) {
    h->F.emplace_back(new pdos_amd::Family);
    std::string err;
    if (!h->F.back()->create(d, k, o, err)) {
      g_fam_err = err;
      delete h;
      return nullptr;
    }
  }
  return h;
}

int pdos_family_load(PdosFamily* h, int count, const double* C, const double* B) {
  std::string err;
  if (!h || count < 1 || !C) { g_fam_err = "bad arguments"; return -1; }
  const int parts = std::min<int>(h->max_engines, (int)h->F.size()) > 1 && count >= kSplitMin ? 2 : 1;
  h->first.clear();
  h->count.clear();
  const int n = h->F[0]->n, m = h->F[0]->m;
  int at = 0;
  for (int p = 0; p < parts; ++p) {
    const int c = p + 1 == parts ? count - at : ((count / parts + 63) & ~63);
    if (!h->F[p]->load(c, C + (size_t)at * n, B ? B + (size_t)at * m : nullptr, err)) {
      g_fam_err = err;
      return -1;
    }
    h->first.push_back(at);
    h->count.push_back(c);
    at += c;
  }
  return 0;
}

int pdos_family_run(PdosFamily* h, const float* warm_dev, long ld, int judge, double* kernel_ms) {
  if (!h || h->count.empty()) { g_fam_err = "nothing loaded"; return -1; }
  const int parts = (int)h->count.size();
  std::vector<std::string> errs(parts);
  std::vector<char> ok(parts, 0);
  auto go = [&](int p) {
    const float* w = warm_dev ? warm_dev + (size_t)h->first[p] * ld : nullptr;
    ok[p] = h->F[p]->run(w, ld, judge, errs[p]) ? 1 : 0;
  };
  if (parts == 1 || h->F[0]->timing) {  // per-kernel timing: one engine at a time
    for (int p = 0; p < parts; ++p) go(p);
  } else {
    std::vector<std::thread> th;
    for (int p = 1; p < parts; ++p) th.emplace_back(go, p);
    go(0);
    for (std::thread& t : th) t.join();
  }
  double ms = 0.0;
  for (int p = 0; p < parts; ++p) {
    if (!ok[p]) { g_fam_err = errs[p]; return -1; }
    ms = std::max(ms, (double)h->F[p]->last_ms);
  }
  if (kernel_ms) *kernel_ms = ms;
  return 0;
}

int pdos_family_engines(PdosFamily* h, int n) {
  if (!h) return -1;
  if (n > 0) h->max_engines = std::min<int>(n, (int)h->F.size());
  return std::min<int>(h->max_engines, (int)h->F.size());
}

int pdos_family_scales(PdosFamily* h, double* lam, double* pscale, double* dscale) {
  if (!h || h->count.empty()) { g_fam_err = "nothing loaded"; return -1; }
  for (size_t p = 0; p < h->count.size(); ++p) {
    pdos_amd::Family& F = *h->F[p];
    const long f = h->first[p], c = h->count[p];
    bool ok = true;
    if (lam) ok = ok && hipMemcpy(lam + f, F.base_lam, c * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (pscale) ok = ok && hipMemcpy(pscale + f, F.base_ps, c * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (dscale) ok = ok && hipMemcpy(dscale + f, F.base_ds, c * 8, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) { g_fam_err = "hipMemcpy failed"; return -1; }
  }
  return 0;
}

int pdos_family_info(PdosFamily* h, long* out) {
  if (!h) return -1;
  const pdos_amd::Family& F = *h->F[0];
  out[0] = F.n;
  out[1] = F.m;
  out[2] = F.np;
  out[3] = 0;
  for (size_t p = 0; p < h->count.size(); ++p) out[3] += h->F[p]->compactions;
  out[4] = F.dense ? 1 : 0;
  out[5] = F.fp64 ? 1 : 0;
  out[6] = F.nprod;
  out[7] = F.hybrid() ? F.nh : 0;
  out[8] = h->count.empty() ? 1 : (long)h->count.size();
  return 0;
}

int pdos_family_fetch(PdosFamily* h, double* X, double* Y, double* S, long* iters, long* status,
                      long* fallback, double* resid) {
  std::string err;
  if (!h) return -1;
  const long n = h->F[0]->n, m = h->F[0]->m;
  for (size_t p = 0; p < h->count.size(); ++p) {
    const long f = h->first[p];
    if (!h->F[p]->fetch(X ? X + f * n : nullptr, Y ? Y + f * m : nullptr, S ? S + f * m : nullptr,
                        iters ? iters + f : nullptr, status ? status + f : nullptr,
                        fallback ? fallback + f : nullptr, resid ? resid + f * 5 : nullptr, err)) {
      g_fam_err = err;
      return -1;
    }
  }
  return 0;
}

void pdos_family_free(PdosFamily* h) { delete h; }

int pdos_family_timing(PdosFamily* h, int enable, double* ms, long* calls, double* cols) {
  if (!h) return -1;
  for (int k = 0; k < pdos_amd::Family::K_KINDS; ++k) {
    if (ms) ms[k] = 0;
    if (calls) calls[k] = 0;
    if (cols) cols[k] = 0;
  }
  for (auto& Fp : h->F) {
    pdos_amd::Family& F = *Fp;
    for (int k = 0; k < pdos_amd::Family::K_KINDS; ++k) {
      if (ms) ms[k] += F.kind_ms[k];
      if (calls) calls[k] += F.kind_calls[k];
      if (cols) cols[k] += F.kind_cols[k];
      F.kind_ms[k] = 0;
      F.kind_calls[k] = 0;
      F.kind_cols[k] = 0;
    }
    F.timing = enable != 0;
  }
  return 0;
}

}  // extern "C"
