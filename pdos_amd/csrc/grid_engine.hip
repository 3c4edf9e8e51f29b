// grid_engine.hip -- one large cone program on the whole GPU (both plugins).
//
// Each phase of the ADMM iteration is one kernel over the whole GPU, except
// the direct plugin's triangular sweeps (below).
//
// SpMV (CSR-stream): the host cuts the rows into blocks of nearly equal
// nonzero count (at most kCapMax; a longer row is a block of its own), sized
// so that the blocks fill the resident workgroup slots a whole number of
// times.  A workgroup loads all entries of its block at once (up to
// kCapMax / 512 independent, coalesced loads per thread), forms the products
// val[p] * x[idx[p]] in LDS, then sums each row from LDS.  Every launch
// therefore streams A at full memory-level parallelism whatever the row
// lengths.
//
// Reductions: a producing kernel writes one partial per workgroup (plain
// stores); every workgroup of the consuming kernel sums those partials itself
// in a fixed order, so all of them agree bit for bit without fences, atomics
// or an extra launch, and runs are reproducible.  Workgroup 0 of a consumer
// also records the scalar state (rho, beta, flags, residuals) for later
// kernels, which read it after the kernel boundary.
//
// One ADMM iteration (pdos.c:94-120) is a fixed launch sequence, captured
// once into a hipGraph and replayed; kernels return at once when the solve
// has stopped or CG has finished, so the stop iteration is exact.  SpMVs per
// iteration with k CG steps: k+1 with A and k+2 with A' (the reference runs
// k+3 and k+2: the A x row sums of s~ = b - A x are reused by the next
// residual check and CG start, exactly, since x does not change in between).
//
// Direct plugin (direct/private.c:16-56, ldl.c:326-428): the permuted
// right-hand side, the forward sweep with L, the backward sweep with D and
// L', and the un-permute are launches of their own.  A sweep follows the
// host's level sets (rows of a level are independent): its rows are stored
// in level order with their entries contiguous, so a run of consecutive
// levels is one contiguous range.  Runs of narrow levels go to one
// workgroup of 1,024 threads that stages each chunk of levels -- entries,
// row ids, row pointers, level pointers, D -- in LDS with coalesced loads
// and keeps z in LDS when it fits, so a level costs LDS accesses and one
// barrier, not global round trips; levels of many rows (a leaf level of a
// large problem) are one launch over the whole GPU each.  The sweeps run in
// binary64 in both precision modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "engine.hpp"

namespace pdos_amd {

namespace {

constexpr int GT = 512;          // threads per workgroup
constexpr int GW = GT / 64;
constexpr int kCapMax = 5000;    // nonzeros per SpMV row block (LDS: 4 blocks per CU in FP64)
constexpr int PER = (kCapMax + GT - 1) / GT;
constexpr int RPT = 2;           // rows per thread: a row block holds at most GT * RPT rows
constexpr int kVecBlocks = 1024; // most workgroups of a vector pass

// SpMV kernels keep 8 waves per SIMD (4 workgroups per CU) resident.
#define SPMV_OCC __attribute__((amdgpu_waves_per_eu(8, 8)))

template <typename T> struct PairOf;
template <> struct PairOf<double> { using type = double2; };
template <> struct PairOf<float> { using type = float2; };
template <typename T> using Pair = typename PairOf<T>::type;
__device__ __forceinline__ double2 make_pair2(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ float2 make_pair2(float a, float b) { return make_float2(a, b); }

// One triangular sweep in level order: position q solves z[row[q]] from the
// entries [ptr[q], ptr[q+1]) (indices into z, binary64 values) and, for the
// backward sweep, divides by dq[q] (D of L D L' at that row) first.
struct SweepDev {
  const int* row;
  const int* ptr;
  const int* idx;
  const double* val;
  const double* dq;
  const int* lptr;   // level -> first position
  const int4* chunk; // staged chunks {lv0, lv1, e0, e1} of runs of narrow levels
};

struct GridScal {
  double rho[2], beta;  // rho of CG step k in rho[k & 1]
  double pres, dres, pobj, dobj, eta;
  int done, status, iters, active, cg_total, pad;
};

template <typename T>
struct GArgs {
  int n, m, f, lend, nsoc, cg_max;
  int nb_a, nb_at, nb_m, nb_n;     // row blocks of A, of A'; vector-pass blocks over m, over n
  double eps, relax, tol_sq, lambda, pscale, dscale;
  const int* rp; const int* ri; const T* rx; const int4* rblk;   // A rows
  const int* cp; const int* ci; const T* cx; const int4* cblk;   // A' rows
  const double* b; const double* c; const double* D; const double* E;
  const int* soc_ptr;
  T* x; T* s; T* y; T* ax; T* st; T* t; T* u;
  Pair<T>* pq;     // [n] CG direction p (.x) and residual q (.y), side by side: one gather for both
  double* p_res;   // [nb_at + nb_m][4]  residual check
  double* p_rho;   // [nb_at]            |q|^2 at the CG start
  double* p_pu;    // [nb_at]            p'u
  double* p_q;     // [nb_n]             |q|^2 after the update
  GridScal* sc;
  double* hist;    // VERBOSE: residuals of checks 0, 10, 20, ... (5 each), or null
  int hist_rows;
  // direct plugin
  int direct;      // 1: direct plugin (no CG start argument t)
  int N, nb_N;     // n + m; vector-pass blocks over N
  const int* perm; // perm[k] = KKT index of pivot k
  double* z;       // [N] right-hand side and solution, pivot order
  SweepDev fw, bw; // forward (L) and backward (D, L') sweeps
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sum of NV values; every thread receives the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w * 4 + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < GW; ++j) a += red[j * 4 + k];
    v[k] = a;
  }
}

// Sums `count` partials of NV components each, in a fixed order.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int count, double (&v)[NV],
                                             double* red) {
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int i = threadIdx.x; i < count; i += GT) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] += part[(size_t)i * NV + k];
  }
  block_sum<NV>(v, red);
}

// One CSR-stream row block {r0, r1, p0, p1}.  load() issues every global load
// of the block that does not depend on the vector being multiplied (all of a
// thread's matrix entries and row pointers at once, so they are in flight
// together); finish() forms the products in LDS and sums each row from LDS
// in order.  A block holding a single row longer than kCapMax is summed by
// the whole workgroup instead.
template <typename T>
struct RowBlock {
  int4 d;
  int ci[PER];
  T cv[PER];
  int q0[RPT], q1[RPT];

  __device__ __forceinline__ void load(const int4* __restrict__ blk, const int* __restrict__ ptr,
                                       const int* __restrict__ idx, const T* __restrict__ val) {
    d = blk[blockIdx.x];
    const int len = d.w - d.z;
    if (len > kCapMax) return;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = (int)threadIdx.x + u * GT;
      if (e < len) { ci[u] = idx[d.z + e]; cv[u] = val[d.z + e]; }
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int r = d.x + (int)threadIdx.x + k * GT;
      if (r < d.y) { q0[k] = ptr[r]; q1[k] = ptr[r + 1]; }
    }
  }

  template <typename Gather, typename Emit>
  __device__ __forceinline__ void finish(const int* __restrict__ idx, const T* __restrict__ val, Gather gx,
                                         Emit emit, T* prod, double* red) {
    const int len = d.w - d.z;
    if (len <= kCapMax) {
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int e = (int)threadIdx.x + u * GT;
        if (e < len) prod[e] = cv[u] * gx(ci[u]);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int r = d.x + (int)threadIdx.x + k * GT;
        if (r < d.y) {
          T acc = T(0);
          for (int e = q0[k] - d.z; e < q1[k] - d.z; ++e) acc += prod[e];
          emit(r, acc);
        }
      }
    } else {
      double acc[1] = {0.0};
      for (int e = d.z + (int)threadIdx.x; e < d.w; e += GT) acc[0] += double(val[e] * gx(idx[e]));
      block_sum<1>(acc, red);
      if (threadIdx.x == 0) emit(d.x, T(acc[0]));
    }
  }
};

template <typename T, typename Gather, typename Emit>
__device__ __forceinline__ void stream_block(const int4* __restrict__ blk, const int* __restrict__ ptr,
                                             const int* __restrict__ idx, const T* __restrict__ val, Gather gx,
                                             Emit emit, T* prod, double* red) {
  RowBlock<T> rb;
  rb.load(blk, ptr, idx, val);
  rb.finish(idx, val, gx, emit, prod, red);
}

template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_ax(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  stream_block<T>(a.rblk, a.rp, a.ri, a.rx, [&](int c) { return a.x[c]; }, [&](int j, T v) { a.ax[j] = v; },
                  prod, red);
}

// Stop-check partials (linAlg.h:158-201): workgroups [0, nb_at) run A'y,
// the rest the m-row vector pass (which also forms the CG start argument t).
template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_resid(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  if (a.sc->done) return;
  double v[4] = {0.0, 0.0, 0.0, 0.0};  // |pres|^2, |dres|^2, c'x, b'y
  const int bid = blockIdx.x;
  if (bid < a.nb_at) {
    stream_block<T>(a.cblk, a.cp, a.ci, a.cx, [&](int c) { return a.y[c]; }, [&](int i, T w) {
      const T ci = T(a.c[i]);
      const T r = (ci + w) / T(a.E[i]);
      v[1] += double(r) * double(r);
      v[2] += double(ci) * double(a.x[i]);
    }, prod, red);
  } else {
    const T lam = T(a.lambda);
    const int per = (a.m + a.nb_m - 1) / a.nb_m, j0 = (bid - a.nb_at) * per, j1 = min(a.m, j0 + per);
    for (int j = j0 + (int)threadIdx.x; j < j1; j += GT) {
      const T sj = a.s[j], bj = T(a.b[j]), yj = a.y[j], axj = a.ax[j];
      const T r = ((sj - bj) + axj) / T(a.D[j]);
      v[0] += double(r) * double(r);
      v[3] += double(bj) * double(yj);
      if (!a.direct) a.t[j] = ((sj + lam * yj) - bj) + axj;
    }
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) a.p_res[(size_t)bid * 4 + k] = v[k];
}

// The stop test (pdos.c:95-107) from the residual partials: every workgroup
// sums them in the same order and reaches the same decision; workgroup 0
// records the residuals, the iteration count and the VERBOSE history.
template <typename T>
__device__ __forceinline__ bool stop_check(const GArgs<T>& a, double* red) {
  double tot[4];
  sum_partials<4>(a.p_res, a.nb_at + a.nb_m, tot, red);
  const double pres = sqrt(tot[0]), dres = sqrt(tot[1]), pobj = tot[2], dobj = -tot[3];
  const double eta = fabs(pobj - dobj);
  const bool stop = pres < a.eps * a.pscale && dres < a.eps * a.dscale &&
                    eta < a.eps * (1.0 + fabs(pobj) + fabs(dobj));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    GridScal* sc = a.sc;
    sc->pres = pres; sc->dres = dres; sc->pobj = pobj; sc->dobj = dobj; sc->eta = eta;
    if (stop) {
      sc->done = 1; sc->status = 0;
    } else {
      const int i = sc->iters;
      if (a.hist && i % 10 == 0 && i / 10 < a.hist_rows) {
        double* h = a.hist + (size_t)(i / 10) * 5;
        h[0] = pres; h[1] = dres; h[2] = pobj; h[3] = dobj; h[4] = eta;
      }
      sc->iters += 1;
    }
  }
  return stop;
}

// The stop test, then the CG start q = -lam c - A't, p = 0 and |q|^2
// partials.
template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_cginit(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  if (a.sc->done) return;
  if (stop_check(a, red)) return;
  const T lam = T(a.lambda);
  double v[1] = {0.0};
  stream_block<T>(a.cblk, a.cp, a.ci, a.cx, [&](int c) { return a.t[c]; }, [&](int i, T w) {
    const T qi = (-lam) * T(a.c[i]) - w;
    a.pq[i] = make_pair2(T(0), qi);
    v[0] += double(qi) * double(qi);
  }, prod, red);
  block_sum<1>(v, red);
  if (threadIdx.x == 0) a.p_rho[blockIdx.x] = v[0];
}

// CG step k, part 1: decide whether CG continues (from the previous step's
// |q|^2 partials), then t = A (beta p + q).  Workgroup 0 stores rho in
// rho[k & 1] while the others still read rho[(k - 1) & 1].
template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_cg_a(GArgs<T> a, int k) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  GridScal* sc = a.sc;
  if (sc->done || (k > 0 && !sc->active)) return;
  RowBlock<T> rb;
  rb.load(a.rblk, a.rp, a.ri, a.rx);
  double rho[1], beta;
  bool active;
  if (k == 0) {
    sum_partials<1>(a.p_rho, a.nb_at, rho, red);
    beta = 0.0;
    active = rho[0] > a.tol_sq && a.cg_max > 0;
  } else {
    sum_partials<1>(a.p_q, a.nb_n, rho, red);
    active = !(rho[0] < a.tol_sq);  // indirect/private.c:86-89
    beta = rho[0] / sc->rho[(k - 1) & 1];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->active = active ? 1 : 0;
    if (active) { sc->rho[k & 1] = rho[0]; sc->beta = beta; }
  }
  if (!active) return;
  const T bt = T(beta);
  rb.finish(a.ri, a.rx, [&](int c) { const Pair<T> v = a.pq[c]; return bt * v.x + v.y; },
            [&](int j, T w) { a.t[j] = w; }, prod, red);
}

// CG step k, part 2: p = beta p + q, u = p + A't, p'u partials.
template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_cg_at(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  const GridScal* sc = a.sc;
  if (sc->done || !sc->active) return;
  const T beta = T(sc->beta);
  double v[1] = {0.0};
  stream_block<T>(a.cblk, a.cp, a.ci, a.cx, [&](int c) { return a.t[c]; }, [&](int i, T w) {
    const Pair<T> pq = a.pq[i];
    const T pi = beta * pq.x + pq.y;
    a.pq[i] = make_pair2(pi, pq.y);
    const T ui = pi + w;
    a.u[i] = ui;
    v[0] += double(pi) * double(ui);
  }, prod, red);
  block_sum<1>(v, red);
  if (threadIdx.x == 0) a.p_pu[blockIdx.x] = v[0];
}

// CG step k, part 3: alpha = rho / p'u, x += alpha p, q -= alpha u, |q|^2
// partials.  One element per thread when n <= GT * kVecBlocks; the vector
// loads are issued before the partials are summed.
template <typename T>
__global__ __launch_bounds__(GT) void k_cg_upd(GArgs<T> a, int k) {
  __shared__ double red[GW * 4];
  GridScal* sc = a.sc;
  if (sc->done || !sc->active) return;
  const int per = (a.n + a.nb_n - 1) / a.nb_n, i0 = blockIdx.x * per, i1 = min(a.n, i0 + per);
  const int i = i0 + (int)threadIdx.x;
  T xi = T(0), pi = T(0), qi = T(0), ui = T(0);
  if (i < i1) {
    const Pair<T> v = a.pq[i];
    xi = a.x[i]; pi = v.x; qi = v.y; ui = a.u[i];
  }
  double pu[1];
  sum_partials<1>(a.p_pu, a.nb_at, pu, red);
  const T alpha = T(sc->rho[k & 1] / pu[0]);
  double v[1] = {0.0};
  if (i < i1) {
    a.x[i] = xi + alpha * pi;
    qi = qi - alpha * ui;
    a.pq[i] = make_pair2(pi, qi);
    v[0] += double(qi) * double(qi);
  }
  for (int j = i + GT; j < i1; j += GT) {
    const Pair<T> w = a.pq[j];
    a.x[j] += alpha * w.x;
    const T qj = w.y - alpha * a.u[j];
    a.pq[j] = make_pair2(w.x, qj);
    v[0] += double(qj) * double(qj);
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) {
    a.p_q[blockIdx.x] = v[0];
    if (blockIdx.x == 0) sc->cg_total += 1;
  }
}

// s~ = b - A x, relax, zero / nonneg cones and their dual update; SOC rows
// keep s~ and z = s~ - lambda y for k_cone (pdos.c:232-249, 296-303).
// Direct plugin: s~ comes from the KKT solve (k_post) instead of b - A x;
// A x is still formed for the next residual check.
template <typename T, bool DIRECT>
__global__ __launch_bounds__(GT) SPMV_OCC void k_fin(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  if (a.sc->done) return;
  const T lam = T(a.lambda), ra = T(a.relax), rb = T(1.0) - T(a.relax);
  stream_block<T>(a.rblk, a.rp, a.ri, a.rx, [&](int c) { return a.x[c]; }, [&](int j, T w) {
    a.ax[j] = w;
    T stj = DIRECT ? a.st[j] : T(a.b[j]) - w;
    stj = ra * stj + rb * a.s[j];
    const T z = stj - lam * a.y[j];
    if (j < a.f) {
      a.s[j] = T(0);
      a.y[j] += (T(0) - stj) / lam;
    } else if (j < a.lend) {
      const T sj = z < T(0) ? T(0) : z;
      a.s[j] = sj;
      a.y[j] += (sj - stj) / lam;
    } else {
      a.st[j] = stj;
      a.s[j] = z;
    }
  }, prod, red);
}

// One wave per second-order cone (cones.c:26-47) + dual update.
template <typename T>
__global__ __launch_bounds__(GT) void k_cone(GArgs<T> a) {
  if (a.sc->done) return;
  const T lam = T(a.lambda);
  const int lane = threadIdx.x & 63;
  for (int cn = blockIdx.x * GW + (threadIdx.x >> 6); cn < a.nsoc; cn += gridDim.x * GW) {
    const int r0 = a.soc_ptr[cn], r1 = a.soc_ptr[cn + 1];
    double a2 = 0.0;
    for (int j = r0 + 1 + lane; j < r1; j += 64) a2 += double(a.s[j]) * double(a.s[j]);
    a2 = wave_sum(a2);
    const T ns = T(sqrt(a2));
    const T t0 = a.s[r0];
    const int mode = ns <= t0 ? 0 : (ns <= -t0 ? 1 : 2);
    const T al = (ns + t0) / T(2);
    for (int j = r0 + lane; j < r1; j += 64) {
      const T zj = a.s[j];
      const T sj = mode == 0 ? zj : (mode == 1 ? T(0) : (j == r0 ? al : al * zj / ns));
      a.s[j] = sj;
      a.y[j] += (sj - a.st[j]) / lam;
    }
  }
}

// ---------------------------------------------------------------- direct

// The stop test, then the permuted KKT right-hand side (direct/private.c:
// 16-31, ldl.c:395-405): z[k] = v[perm[k]], v = [lam c - x ; b - (s + lam y)].
template <typename T>
__global__ __launch_bounds__(GT) void k_rhs(GArgs<T> a) {
  __shared__ double red[GW * 4];
  if (a.sc->done) return;
  if (stop_check(a, red)) return;
  const double lam = a.lambda;
  const int per = (a.N + a.nb_N - 1) / a.nb_N, k0 = blockIdx.x * per, k1 = min(a.N, k0 + per);
  for (int k = k0 + (int)threadIdx.x; k < k1; k += GT) {
    const int src = a.perm[k];
    double v;
    if (src < a.n) {
      v = lam * a.c[src] - double(a.x[src]);
    } else {
      const int j = src - a.n;
      v = a.b[j] - (double(a.s[j]) + lam * double(a.y[j]));
    }
    a.z[k] = v;
  }
}

// Un-permute (ldl.c:412-428) and direct/private.c:48-55: x = z[0:n],
// s~ = z[n:] + (s + lam y).
template <typename T>
__global__ __launch_bounds__(GT) void k_post(GArgs<T> a) {
  if (a.sc->done) return;
  const double lam = a.lambda;
  const int per = (a.N + a.nb_N - 1) / a.nb_N, k0 = blockIdx.x * per, k1 = min(a.N, k0 + per);
  for (int k = k0 + (int)threadIdx.x; k < k1; k += GT) {
    const int dst = a.perm[k];
    const double v = a.z[k];
    if (dst < a.n) {
      a.x[dst] = T(v);
    } else {
      const int j = dst - a.n;
      a.st[j] = T(v + (double(a.s[j]) + lam * double(a.y[j])));
    }
  }
}

constexpr int ST = 1024;        // threads of the staged sweep workgroup
constexpr int SW = ST / 64;
constexpr int kRunRows = 1024;  // rows (positions) a staged chunk holds
constexpr int kRunLevels = 1024;  // levels a staged chunk holds

// Solves the positions [q0, q1) of one level.  Rows get a wave each when the
// level is short or its rows are long (lanes stride the row, then a wave
// sum), else a thread each (entries in storage order).  Acc supplies row(q),
// ptr(q), idx(e), val(e), dq(q) (staged or global) and Z the vector.
template <bool BACK, typename Acc, typename Z>
__device__ __forceinline__ void sweep_level(int q0, int q1, const Acc& ac, Z& zv) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nrows = q1 - q0;
  if (nrows <= 0) return;
  const bool wave_rows = nrows <= SW || (ac.ptr(q1) - ac.ptr(q0)) >= 32 * nrows;
  if (wave_rows) {
    for (int q = q0 + w; q < q1; q += SW) {
      const int e0 = ac.ptr(q), e1 = ac.ptr(q + 1);
      double acc = 0.0;
      for (int e = e0 + lane; e < e1; e += 64) acc += ac.val(e) * zv.get(ac.idx(e));
      acc = wave_sum(acc);
      if (lane == 0) {
        const int i = ac.row(q);
        zv.set(i, (BACK ? zv.get(i) / ac.dq(q) : zv.get(i)) - acc);
      }
    }
  } else {
    for (int q = q0 + tid; q < q1; q += ST) {
      const int i = ac.row(q), e0 = ac.ptr(q), e1 = ac.ptr(q + 1);
      double v = BACK ? zv.get(i) / ac.dq(q) : zv.get(i);
      for (int e = e0; e < e1; ++e) v -= ac.val(e) * zv.get(ac.idx(e));
      zv.set(i, v);
    }
  }
}

struct GlobalAcc {  // a sweep read from global memory
  const SweepDev* s;
  __device__ int row(int q) const { return s->row[q]; }
  __device__ int ptr(int q) const { return s->ptr[q]; }
  __device__ int idx(int e) const { return s->idx[e]; }
  __device__ double val(int e) const { return s->val[e]; }
  __device__ double dq(int q) const { return s->dq[q]; }
};

struct StagedAcc {  // a chunk staged in LDS (positions and entries rebased)
  const int* srow; const int* sptr; const int* sidx; const double* sval; const double* sdq;
  int q0, e0;
  __device__ int row(int q) const { return srow[q - q0]; }
  __device__ int ptr(int q) const { return sptr[q - q0] + e0; }
  __device__ int idx(int e) const { return sidx[e - e0]; }
  __device__ double val(int e) const { return sval[e - e0]; }
  __device__ double dq(int q) const { return sdq[q - q0]; }
};

struct ZGlobal {
  double* z;
  __device__ double get(int i) const { return z[i]; }
  __device__ void set(int i, double v) const { z[i] = v; }
};
struct ZShared {
  double* z;
  __device__ double get(int i) const { return z[i]; }
  __device__ void set(int i, double v) const { z[i] = v; }
};

// Runs the staged chunks [c0, c1) of one sweep in one workgroup.  Dynamic
// LDS: [z (N doubles, when ZLDS)] [entries: cap doubles, cap ints]
// [rows: kRunRows ints, kRunRows + 1 ptrs, kRunRows doubles] [kRunLevels + 1
// level pointers].  A chunk whose level alone exceeds cap entries or
// kRunRows rows is read from global memory instead.
template <typename T, bool BACK, bool ZLDS>
__global__ __launch_bounds__(ST) void k_sweep_run(GArgs<T> a, int c0, int c1, int cap) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  if (a.sc->done) return;
  const SweepDev& sw = BACK ? a.bw : a.fw;
  const int tid = threadIdx.x;
  unsigned char* ptrb = smem;
  double* zs = nullptr;
  if (ZLDS) {
    zs = reinterpret_cast<double*>(ptrb);
    ptrb += (((size_t)a.N * 8) + 15) & ~(size_t)15;
  }
  double* sval = reinterpret_cast<double*>(ptrb);
  ptrb += (size_t)cap * 8;
  double* sdq = reinterpret_cast<double*>(ptrb);
  ptrb += (size_t)kRunRows * 8;
  int* sidx = reinterpret_cast<int*>(ptrb);
  ptrb += (((size_t)cap * 4) + 15) & ~(size_t)15;
  int* srow = reinterpret_cast<int*>(ptrb);
  ptrb += (size_t)kRunRows * 4;
  int* sptr = reinterpret_cast<int*>(ptrb);
  ptrb += (size_t)(kRunRows + 4) * 4;
  int* slp = reinterpret_cast<int*>(ptrb);
  if (ZLDS)
    for (int i = tid; i < a.N; i += ST) zs[i] = a.z[i];
  auto run_levels = [&](auto& zv) {
    for (int c = c0; c < c1; ++c) {
      const int4 ch = sw.chunk[c];
      const int lv0 = ch.x, lv1 = ch.y, e0 = ch.z, e1 = ch.w;
      const int p0 = sw.lptr[lv0], p1 = sw.lptr[lv1];
      const bool staged = e1 - e0 <= cap && p1 - p0 <= kRunRows && lv1 - lv0 <= kRunLevels;
      __syncthreads();  // the previous chunk's staged data and z writes are done
      if (staged) {
        for (int e = e0 + tid; e < e1; e += ST) { sval[e - e0] = sw.val[e]; sidx[e - e0] = sw.idx[e]; }
        for (int q = p0 + tid; q <= p1; q += ST) {
          sptr[q - p0] = sw.ptr[q] - e0;
          if (q < p1) {
            srow[q - p0] = sw.row[q];
            if (BACK) sdq[q - p0] = sw.dq[q];
          }
        }
        for (int l = lv0 + tid; l <= lv1; l += ST) slp[l - lv0] = sw.lptr[l];
        __syncthreads();
        const StagedAcc ac{srow, sptr, sidx, sval, sdq, p0, e0};
        for (int lv = lv0; lv < lv1; ++lv) {
          sweep_level<BACK>(slp[lv - lv0], slp[lv - lv0 + 1], ac, zv);
          __syncthreads();
        }
      } else {
        const GlobalAcc ac{&sw};
        for (int lv = lv0; lv < lv1; ++lv) {
          sweep_level<BACK>(sw.lptr[lv], sw.lptr[lv + 1], ac, zv);
          __syncthreads();
        }
      }
    }
  };
  if (ZLDS) {
    ZShared zv{zs};
    run_levels(zv);
    __syncthreads();
    for (int i = tid; i < a.N; i += ST) a.z[i] = zs[i];
  } else {
    ZGlobal zv{a.z};
    run_levels(zv);
  }
}

// One wide level over the whole GPU: a thread per row.
template <typename T, bool BACK>
__global__ __launch_bounds__(GT) void k_sweep_wide(GArgs<T> a, int lv) {
  if (a.sc->done) return;
  const SweepDev& sw = BACK ? a.bw : a.fw;
  const int q = sw.lptr[lv] + blockIdx.x * GT + (int)threadIdx.x;
  if (q >= sw.lptr[lv + 1]) return;
  const int i = sw.row[q], e0 = sw.ptr[q], e1 = sw.ptr[q + 1];
  double v = BACK ? a.z[i] / sw.dq[q] : a.z[i];
  for (int e = e0; e < e1; ++e) v -= sw.val[e] * a.z[sw.idx[e]];
  a.z[i] = v;
}

// Host schedule of one sweep: what each launch covers.
struct SweepSeg {
  int wide;      // 1: one wide level `lv` (k_sweep_wide); 0: chunks [c0, c1) (k_sweep_run)
  int lv, c0, c1;
  int rows;      // wide: rows of the level
};

// Host schedule of both sweeps plus what the launches need.
struct SweepPlan {
  std::vector<SweepSeg> fw, bw;
  bool zlds = false;
  int cap = 0, lds = 0;
};

template <typename T, bool BACK>
void enqueue_sweep(const GArgs<T>& a, const std::vector<SweepSeg>& segs, const SweepPlan& pl, hipStream_t st) {
  for (const SweepSeg& g : segs) {
    if (g.wide) {
      hipLaunchKernelGGL((k_sweep_wide<T, BACK>), dim3((g.rows + GT - 1) / GT), dim3(GT), 0, st, a, g.lv);
    } else if (pl.zlds) {
      hipLaunchKernelGGL((k_sweep_run<T, BACK, true>), dim3(1), dim3(ST), pl.lds, st, a, g.c0, g.c1, pl.cap);
    } else {
      hipLaunchKernelGGL((k_sweep_run<T, BACK, false>), dim3(1), dim3(ST), pl.lds, st, a, g.c0, g.c1, pl.cap);
    }
  }
}

template <typename T>
void enqueue_iteration(const GArgs<T>& a, const SweepPlan& pl, hipStream_t st) {
  const int cone_blocks = std::max(1, std::min(2048, (a.nsoc + GW - 1) / GW));
  hipLaunchKernelGGL(k_resid<T>, dim3(a.nb_at + a.nb_m), dim3(GT), 0, st, a);
  if (a.direct) {
    hipLaunchKernelGGL(k_rhs<T>, dim3(a.nb_N), dim3(GT), 0, st, a);
    enqueue_sweep<T, false>(a, pl.fw, pl, st);
    enqueue_sweep<T, true>(a, pl.bw, pl, st);
    hipLaunchKernelGGL(k_post<T>, dim3(a.nb_N), dim3(GT), 0, st, a);
    hipLaunchKernelGGL((k_fin<T, true>), dim3(a.nb_a), dim3(GT), 0, st, a);
  } else {
    hipLaunchKernelGGL(k_cginit<T>, dim3(a.nb_at), dim3(GT), 0, st, a);
    for (int k = 0; k < a.cg_max; ++k) {
      hipLaunchKernelGGL(k_cg_a<T>, dim3(a.nb_a), dim3(GT), 0, st, a, k);
      hipLaunchKernelGGL(k_cg_at<T>, dim3(a.nb_at), dim3(GT), 0, st, a);
      hipLaunchKernelGGL(k_cg_upd<T>, dim3(a.nb_n), dim3(GT), 0, st, a, k);
    }
    hipLaunchKernelGGL((k_fin<T, false>), dim3(a.nb_a), dim3(GT), 0, st, a);
  }
  hipLaunchKernelGGL(k_cone<T>, dim3(cone_blocks), dim3(GT), 0, st, a);
}

// One sweep's arrays in level order.  Forward: row i of L (CSR) at each
// position; backward: column j of L (a row of L') and D[j].
struct SweepHost {
  std::vector<int> row, ptr, idx, lptr;
  std::vector<double> val, dq;
  std::vector<int4> chunk;
};

// Levels of at least `wide_rows` rows become wide launches; each run of
// narrower levels is cut into chunks of whole levels holding at most `cap`
// entries, kRunRows rows and kRunLevels levels (a level alone beyond that is
// a chunk of its own, read from global memory by the run kernel).
void build_sweep(const Factor& F, bool back, int wide_rows, int cap, SweepHost& h, std::vector<SweepSeg>& segs) {
  const int N = F.N;
  const std::vector<int>& lp = back ? F.bl_ptr : F.fl_ptr;
  const std::vector<int>& lr = back ? F.bl_rows : F.fl_rows;
  const std::vector<int>& P = back ? F.Lp : F.Rp;
  const std::vector<int>& I = back ? F.Li : F.Ri;
  const std::vector<double>& V = back ? F.Lx : F.Rx;
  const int nlev = (int)lp.size() - 1;
  h.row.assign(lr.begin(), lr.end());
  h.lptr.assign(lp.begin(), lp.end());
  h.ptr.assign(N + 1, 0);
  for (int q = 0; q < N; ++q) h.ptr[q + 1] = h.ptr[q] + (P[h.row[q] + 1] - P[h.row[q]]);
  h.idx.resize(h.ptr[N]);
  h.val.resize(h.ptr[N]);
  for (int q = 0; q < N; ++q) {
    const int r = h.row[q];
    std::copy(I.begin() + P[r], I.begin() + P[r + 1], h.idx.begin() + h.ptr[q]);
    std::copy(V.begin() + P[r], V.begin() + P[r + 1], h.val.begin() + h.ptr[q]);
  }
  if (back) {
    h.dq.resize(N);
    for (int q = 0; q < N; ++q) h.dq[q] = F.Ld[h.row[q]];
  }
  auto rows_of = [&](int lv) { return h.lptr[lv + 1] - h.lptr[lv]; };
  auto ents = [&](int l0, int l1) { return h.ptr[h.lptr[l1]] - h.ptr[h.lptr[l0]]; };
  h.chunk.clear();
  segs.clear();
  int lv = 0;
  while (lv < nlev) {
    if (rows_of(lv) >= wide_rows) {
      segs.push_back(SweepSeg{1, lv, 0, 0, rows_of(lv)});
      ++lv;
      continue;
    }
    SweepSeg g{0, 0, (int)h.chunk.size(), 0, 0};
    while (lv < nlev && rows_of(lv) < wide_rows) {
      int l1 = lv + 1;
      while (l1 < nlev && rows_of(l1) < wide_rows && ents(lv, l1 + 1) <= cap &&
             h.lptr[l1 + 1] - h.lptr[lv] <= kRunRows && l1 + 1 - lv <= kRunLevels)
        ++l1;
      h.chunk.push_back(int4{lv, l1, h.ptr[h.lptr[lv]], h.ptr[h.lptr[l1]]});
      lv = l1;
    }
    g.c1 = (int)h.chunk.size();
    segs.push_back(g);
  }
}

// Row blocks {r0, r1, p0, p1} for CSR-stream.  The nonzero cap is chosen so
// that the blocks fill `slots` resident workgroups a whole number of times
// (a partly filled last round would cost as long as a full one); at most
// GT * RPT rows per block; a row longer than kCapMax is a block of its own.
std::vector<int4> row_blocks(const std::vector<int>& ptr, int slots) {
  const int nrows = (int)ptr.size() - 1;
  const long nnz = nrows > 0 ? ptr[nrows] : 0;
  int maxrow = 0;
  for (int r = 0; r < nrows; ++r) maxrow = std::max(maxrow, std::min(ptr[r + 1] - ptr[r], kCapMax));
  long cap = kCapMax;
  for (long rounds = 1; rounds < 64; ++rounds) {
    const long c = (nnz + rounds * slots - 1) / (rounds * slots) + maxrow;
    if (c <= kCapMax) { cap = std::max<long>(c, std::min<long>(512, kCapMax)); break; }
  }
  std::vector<int4> out;
  int r = 0;
  while (r < nrows) {
    int e = r;
    while (e < nrows && ptr[e + 1] - ptr[r] <= cap && e - r < GT * RPT) ++e;
    if (e == r) e = r + 1;  // a row longer than the cap
    out.push_back(int4{r, e, ptr[r], ptr[e]});
    r = e;
  }
  if (out.empty()) out.push_back(int4{0, 0, 0, 0});
  return out;
}

template <typename T>
size_t put_as(Packer& P, const std::vector<double>& v) {
  std::vector<T> w(v.begin(), v.end());
  return P.put(w.data(), w.size() * sizeof(T));
}

template <typename T>
class GridImpl : public GridSolver {
 public:
  ~GridImpl() override {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (own) {
      if (dbuf) (void)hipFree(dbuf);
      if (hbuf) (void)hipHostFree(hbuf);
      if (stream) (void)hipStreamDestroy(stream);
    }
  }

  bool prepare(const Setup& S, const Params& par, const Factor* F, int device, bool borrow, std::string& err) {
    Ctx& C = ctx();
    if (!C.ensure(device, err)) return false;
    own = !borrow;
    const double t0 = now_ms();
    n = S.n; m = S.m;
    const int N = n + m;
    int occ = 0;  // resident SpMV workgroups per CU (4 by design)
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_cg_a<T>, GT, 0) != hipSuccess || occ < 1) occ = 4;
    const int slots = occ * C.num_cu;
    const std::vector<int4> rb = row_blocks(S.rp, slots), cb = row_blocks(S.cp, slots);
    const int nb_m = std::max(1, std::min(kVecBlocks, (m + GT - 1) / GT));
    const int nb_n = std::max(1, std::min(kVecBlocks, (n + GT - 1) / GT));
    Packer P;
    const size_t o_rp = P.put(S.rp.data(), S.rp.size() * 4), o_ri = P.put(S.ri.data(), S.ri.size() * 4);
    const size_t o_rx = put_as<T>(P, S.rx), o_rb = P.put(rb.data(), rb.size() * sizeof(int4));
    const size_t o_cp = P.put(S.cp.data(), S.cp.size() * 4), o_ci = P.put(S.ci.data(), S.ci.size() * 4);
    const size_t o_cx = put_as<T>(P, S.cx), o_cb = P.put(cb.data(), cb.size() * sizeof(int4));
    const size_t o_b = P.put(S.b.data(), m * 8), o_c = P.put(S.c.data(), n * 8);
    const size_t o_D = P.put(S.D.data(), m * 8), o_E = P.put(S.E.data(), n * 8);
    const size_t o_soc = P.put(S.soc_ptr.data(), S.soc_ptr.size() * 4);
    const size_t o_x0 = put_as<T>(P, S.x), o_s0 = put_as<T>(P, S.s), o_y0 = put_as<T>(P, S.y);
    const size_t o_x = P.zeros(n * sizeof(T)), o_s = P.zeros(m * sizeof(T)), o_y = P.zeros(m * sizeof(T));
    const size_t o_ax = P.zeros(m * sizeof(T)), o_st = P.zeros(m * sizeof(T));
    const bool direct = F != nullptr;
    const size_t o_t = P.zeros((direct ? 1 : m) * sizeof(T));
    const size_t o_pq = P.zeros((direct ? 1 : n) * sizeof(Pair<T>)), o_u = P.zeros((direct ? 1 : n) * sizeof(T));
    const size_t o_pres = P.zeros((cb.size() + nb_m) * 4 * 8), o_prho = P.zeros(cb.size() * 8);
    const size_t o_ppu = P.zeros(cb.size() * 8), o_pqn = P.zeros(nb_n * 8);
    std::memset(&sc0, 0, sizeof(sc0));
    sc0.status = 1;
    sc0.pres = sc0.dres = sc0.pobj = sc0.dobj = sc0.eta = -1.0;
    const size_t o_sc = P.put(&sc0, sizeof(sc0));
    hrows = par.VERBOSE ? verbose_hist_rows(par.MAX_ITERS) : 0;
    const size_t o_hist = P.zeros(sizeof(double) * 5 * (size_t)std::max(hrows, 1));
    // direct plugin: permutation, z and both sweeps in level order
    SweepHost fw, bw;
    size_t o_perm = 0, o_z = 0, o_f[7] = {0}, o_b7[7] = {0};
    if (direct) {
      const char* wr = std::getenv("PDOS_SWEEP_WIDE_ROWS");
      const int wide_rows = wr && std::atoi(wr) > 0 ? std::atoi(wr) : 8192;
      const int fixed = kRunRows * 8 + kRunRows * 4 + (kRunRows + 4) * 4 + (kRunLevels + 1) * 4 + 64;
      const long zbytes = (((long)N * 8) + 15) & ~15L;
      plan.zlds = zbytes <= 96 * 1024 && zbytes + fixed + 12 * 1024 <= C.max_lds;
      const long left = (long)C.max_lds - fixed - (plan.zlds ? zbytes : 0);
      plan.cap = (int)std::max<long>(256, std::min<long>(1 << 20, (left - 16) / 12) & ~3L);
      plan.lds = (int)((plan.zlds ? zbytes : 0) + (long)plan.cap * 8 + kRunRows * 8 +
                       ((((long)plan.cap * 4) + 15) & ~15L) + kRunRows * 4 + (kRunRows + 4) * 4 +
                       (kRunLevels + 1) * 4);
      if (plan.lds > C.max_lds) { err = "grid engine: sweep workgroup exceeds the LDS"; return false; }
      build_sweep(*F, false, wide_rows, plan.cap, fw, plan.fw);
      build_sweep(*F, true, wide_rows, plan.cap, bw, plan.bw);
      o_perm = P.put(F->P.data(), N * 4);
      o_z = P.zeros((size_t)N * 8);
      SweepHost* hs[2] = {&fw, &bw};
      size_t* os[2] = {o_f, o_b7};
      for (int d = 0; d < 2; ++d) {
        SweepHost& h = *hs[d];
        os[d][0] = P.put(h.row.data(), h.row.size() * 4);
        os[d][1] = P.put(h.ptr.data(), h.ptr.size() * 4);
        os[d][2] = P.put(h.idx.data(), h.idx.size() * 4);
        os[d][3] = P.put(h.val.data(), h.val.size() * 8);
        os[d][4] = P.put(h.dq.data(), h.dq.size() * 8);
        os[d][5] = P.put(h.lptr.data(), h.lptr.size() * 4);
        os[d][6] = P.put(h.chunk.data(), h.chunk.size() * sizeof(int4));
      }
    }
    bytes = P.size();
    const size_t hbytes = std::max<size_t>(bytes, 4096);
    hipError_t e = hipSuccess;
    if (borrow) {
      if (!C.reserve(bytes, hbytes, err)) return false;
      dbuf = C.dbuf;
      hbuf = C.hbuf;
      stream = C.stream;
    } else {
      e = hipMalloc(&dbuf, bytes);
      if (e == hipSuccess) e = hipHostMalloc(&hbuf, hbytes, hipHostMallocDefault);
      if (e == hipSuccess) e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
      if (e != hipSuccess) { err = std::string("grid engine: ") + hipGetErrorString(e); return false; }
    }
    unsigned char* base = static_cast<unsigned char*>(dbuf);
    auto at = [&](size_t off) { return base + off; };
    a.n = n; a.m = m; a.f = S.f; a.lend = S.f + S.l; a.nsoc = S.nsoc();
    a.cg_max = direct ? 0 : (int)std::min<long>(std::max<long>(par.CG_MAX_ITS, 0), 100000);
    a.nb_a = (int)rb.size(); a.nb_at = (int)cb.size(); a.nb_m = nb_m; a.nb_n = nb_n;
    a.eps = par.EPS_ABS; a.relax = par.ALPHA; a.tol_sq = par.CG_TOL * par.CG_TOL;
    a.lambda = S.lambda; a.pscale = S.pscale; a.dscale = S.dscale;
    a.rp = (const int*)at(o_rp); a.ri = (const int*)at(o_ri); a.rx = (const T*)at(o_rx);
    a.rblk = (const int4*)at(o_rb);
    a.cp = (const int*)at(o_cp); a.ci = (const int*)at(o_ci); a.cx = (const T*)at(o_cx);
    a.cblk = (const int4*)at(o_cb);
    a.b = (const double*)at(o_b); a.c = (const double*)at(o_c);
    a.D = (const double*)at(o_D); a.E = (const double*)at(o_E);
    a.soc_ptr = (const int*)at(o_soc);
    a.x = (T*)at(o_x); a.s = (T*)at(o_s); a.y = (T*)at(o_y);
    x0 = at(o_x0); s0 = at(o_s0); y0 = at(o_y0);
    a.ax = (T*)at(o_ax); a.st = (T*)at(o_st); a.t = (T*)at(o_t);
    a.pq = (Pair<T>*)at(o_pq); a.u = (T*)at(o_u);
    a.p_res = (double*)at(o_pres); a.p_rho = (double*)at(o_prho);
    a.p_pu = (double*)at(o_ppu); a.p_q = (double*)at(o_pqn);
    a.sc = (GridScal*)at(o_sc);
    a.hist = hrows ? (double*)at(o_hist) : nullptr;
    a.hist_rows = hrows;
    a.direct = direct ? 1 : 0;
    a.N = N;
    a.nb_N = std::max(1, std::min(kVecBlocks, (N + GT - 1) / GT));
    a.perm = direct ? (const int*)at(o_perm) : nullptr;
    a.z = direct ? (double*)at(o_z) : nullptr;
    std::memset(&a.fw, 0, sizeof(a.fw));
    std::memset(&a.bw, 0, sizeof(a.bw));
    if (direct) {
      SweepDev* sd[2] = {&a.fw, &a.bw};
      size_t* os[2] = {o_f, o_b7};
      for (int d = 0; d < 2; ++d) {
        sd[d]->row = (const int*)at(os[d][0]);
        sd[d]->ptr = (const int*)at(os[d][1]);
        sd[d]->idx = (const int*)at(os[d][2]);
        sd[d]->val = (const double*)at(os[d][3]);
        sd[d]->dq = d ? (const double*)at(os[d][4]) : nullptr;
        sd[d]->lptr = (const int*)at(os[d][5]);
        sd[d]->chunk = (const int4*)at(os[d][6]);
      }
      const void* fns[4] = {reinterpret_cast<const void*>(k_sweep_run<T, false, true>),
                            reinterpret_cast<const void*>(k_sweep_run<T, true, true>),
                            reinterpret_cast<const void*>(k_sweep_run<T, false, false>),
                            reinterpret_cast<const void*>(k_sweep_run<T, true, false>)};
      for (const void* fn : fns)
        if (e == hipSuccess) e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, plan.lds);
    }
    std::memcpy(hbuf, P.data(), bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dbuf, hbuf, bytes, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipEventCreate(&ev0);
    if (e == hipSuccess) e = hipEventCreate(&ev1);
    if (e != hipSuccess) { err = std::string("upload: ") + hipGetErrorString(e); return false; }
    max_iters = std::max<long>(par.MAX_ITERS, 0);
    if (max_iters > 0) {
      e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        enqueue_iteration<T>(a, plan, stream);
        e = hipStreamEndCapture(stream, &graph);
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) { err = std::string("graph: ") + hipGetErrorString(e); return false; }
    }
    upload_ms = now_ms() - t0;
    return true;
  }

  bool run(std::string& err) override {
    const double t1 = now_ms();
    hipError_t e = hipMemcpyAsync(a.x, x0, n * sizeof(T), hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(a.s, s0, m * sizeof(T), hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(a.y, y0, m * sizeof(T), hipMemcpyDeviceToDevice, stream);
    GridScal* hsc = reinterpret_cast<GridScal*>(hbuf);
    *hsc = sc0;
    if (e == hipSuccess) e = hipMemcpyAsync(a.sc, hsc, sizeof(GridScal), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipEventRecord(ev0, stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_ax<T>, dim3(a.nb_a), dim3(GT), 0, stream, a);
      e = hipGetLastError();
    }
    long launched = 0;
    int chunk = 4;
    while (e == hipSuccess && launched < max_iters) {
      const long k = std::min<long>(chunk, max_iters - launched);
      for (long i = 0; i < k && e == hipSuccess; ++i) e = hipGraphLaunch(exec, stream);
      launched += k;
      if (e == hipSuccess) e = hipMemcpyAsync(hsc, a.sc, sizeof(GridScal), hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess || hsc->done) break;
      chunk = std::min(chunk * 2, 32);
    }
    if (e == hipSuccess) e = hipEventRecord(ev1, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(hsc, a.sc, sizeof(GridScal), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    if (e != hipSuccess) { err = std::string("grid engine: ") + hipGetErrorString(e); return false; }
    scv = *hsc;
    kernel_ms = ms;
    solve_ms = now_ms() - t1;
    return true;
  }

  bool fetch(Setup& S, DevResult& r, std::vector<double>* hist, std::string& err) override {
    std::vector<T> hx(n), hs(m), hy(m);
    hipError_t e = hipMemcpy(hx.data(), a.x, n * sizeof(T), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hs.data(), a.s, m * sizeof(T), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hy.data(), a.y, m * sizeof(T), hipMemcpyDeviceToHost);
    if (e == hipSuccess && hist) {
      hist->assign((size_t)5 * hrows, 0.0);
      if (hrows) e = hipMemcpy(hist->data(), a.hist, sizeof(double) * 5 * hrows, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) { err = std::string("download: ") + hipGetErrorString(e); return false; }
    S.x.resize(n);
    S.s.resize(m);
    S.y.resize(m);
    for (int i = 0; i < n; ++i) S.x[i] = double(hx[i]);
    for (int j = 0; j < m; ++j) { S.s[j] = double(hs[j]); S.y[j] = double(hy[j]); }
    r = DevResult{};
    r.iters = scv.iters;
    r.status = scv.status;
    r.cg_iters = scv.cg_total;
    r.p_res = scv.pres; r.d_res = scv.dres; r.p_obj = scv.pobj; r.d_obj = scv.dobj; r.eta = scv.eta;
    return true;
  }

 private:
  GArgs<T> a;
  SweepPlan plan;
  GridScal sc0, scv;
  int n = 0, m = 0, hrows = 0;
  long max_iters = 0;
  size_t bytes = 0;
  bool own = true;
  void* dbuf = nullptr;
  void* hbuf = nullptr;
  void* x0 = nullptr;
  void* s0 = nullptr;
  void* y0 = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

}  // namespace

bool grid_preferred(const Setup& S, bool fp64, bool direct) {
  const int lds = wg_lds_bytes(S.n, S.m, S.n + S.m, direct, fp64 ? 8 : 4);
  return lds > ctx().max_lds || S.nnz() > 200000 || S.m > 16384;
}

std::unique_ptr<GridSolver> grid_prepare(const Setup& S, const Params& par, bool fp64, const Factor* F, int device,
                                         bool borrow, std::string& err) {
  std::unique_ptr<GridSolver> g;
  bool ok;
  if (fp64) {
    auto* p = new GridImpl<double>();
    g.reset(p);
    ok = p->prepare(S, par, F, device, borrow, err);
  } else {
    auto* p = new GridImpl<float>();
    g.reset(p);
    ok = p->prepare(S, par, F, device, borrow, err);
  }
  if (!ok) g.reset();
  return g;
}

bool solve_grid(Setup& S, const Params& par, bool fp64, const Factor* F, int device, std::vector<DevResult>& res,
                double& upload_ms, double& solve_ms, std::string& err, std::vector<double>* hist) {
  std::unique_ptr<GridSolver> g = grid_prepare(S, par, fp64, F, device, true, err);
  if (!g) return false;
  res.assign(1, DevResult{});
  if (!g->run(err) || !g->fetch(S, res[0], hist, err)) return false;
  upload_ms = g->upload_ms;
  solve_ms = g->solve_ms;
  return true;
}

}  // namespace pdos_amd
