// grid_engine.hip -- one large cone program on the whole GPU (indirect plugin).
//
// Each phase of the ADMM iteration is one kernel over the whole GPU.
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
      a.t[j] = ((sj + lam * yj) - bj) + axj;
    }
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) a.p_res[(size_t)bid * 4 + k] = v[k];
}

// The stop test (every workgroup, same order, same result), then the CG
// start q = -lam c - A't, p = 0 and |q|^2 partials.
template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_cginit(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  if (a.sc->done) return;
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
  if (stop) return;
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
template <typename T>
__global__ __launch_bounds__(GT) SPMV_OCC void k_fin(GArgs<T> a) {
  __shared__ T prod[kCapMax];
  __shared__ double red[GW * 4];
  if (a.sc->done) return;
  const T lam = T(a.lambda), ra = T(a.relax), rb = T(1.0) - T(a.relax);
  stream_block<T>(a.rblk, a.rp, a.ri, a.rx, [&](int c) { return a.x[c]; }, [&](int j, T w) {
    a.ax[j] = w;
    T stj = T(a.b[j]) - w;
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

template <typename T>
void enqueue_iteration(const GArgs<T>& a, hipStream_t st) {
  const int cone_blocks = std::max(1, std::min(2048, (a.nsoc + GW - 1) / GW));
  hipLaunchKernelGGL(k_resid<T>, dim3(a.nb_at + a.nb_m), dim3(GT), 0, st, a);
  hipLaunchKernelGGL(k_cginit<T>, dim3(a.nb_at), dim3(GT), 0, st, a);
  for (int k = 0; k < a.cg_max; ++k) {
    hipLaunchKernelGGL(k_cg_a<T>, dim3(a.nb_a), dim3(GT), 0, st, a, k);
    hipLaunchKernelGGL(k_cg_at<T>, dim3(a.nb_at), dim3(GT), 0, st, a);
    hipLaunchKernelGGL(k_cg_upd<T>, dim3(a.nb_n), dim3(GT), 0, st, a, k);
  }
  hipLaunchKernelGGL(k_fin<T>, dim3(a.nb_a), dim3(GT), 0, st, a);
  hipLaunchKernelGGL(k_cone<T>, dim3(cone_blocks), dim3(GT), 0, st, a);
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
bool run_grid(Setup& S, const Params& par, int device, std::vector<DevResult>& res, double& upload_ms,
              double& solve_ms, std::string& err, std::vector<double>* hist) {
  Ctx& C = ctx();
  if (!C.ensure(device, err)) return false;
  const double t0 = now_ms();
  const int n = S.n, m = S.m;
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
  const size_t o_x = put_as<T>(P, S.x), o_s = put_as<T>(P, S.s), o_y = put_as<T>(P, S.y);
  const size_t o_ax = P.zeros(m * sizeof(T)), o_st = P.zeros(m * sizeof(T)), o_t = P.zeros(m * sizeof(T));
  const size_t o_pq = P.zeros(n * sizeof(Pair<T>)), o_u = P.zeros(n * sizeof(T));
  const size_t o_pres = P.zeros((cb.size() + nb_m) * 4 * 8), o_prho = P.zeros(cb.size() * 8);
  const size_t o_ppu = P.zeros(cb.size() * 8), o_pqn = P.zeros(nb_n * 8);
  GridScal sc0;
  std::memset(&sc0, 0, sizeof(sc0));
  sc0.status = 1;
  sc0.pres = sc0.dres = sc0.pobj = sc0.dobj = sc0.eta = -1.0;
  const size_t o_sc = P.put(&sc0, sizeof(sc0));
  const int hrows = par.VERBOSE ? verbose_hist_rows(par.MAX_ITERS) : 0;
  const size_t o_hist = P.zeros(sizeof(double) * 5 * (size_t)std::max(hrows, 1));
  if (!C.reserve(P.buf.size(), std::max(P.buf.size(), (size_t)(n + 2 * m) * 8 + 4096), err)) return false;
  unsigned char* base = static_cast<unsigned char*>(C.dbuf);
  auto at = [&](size_t off) { return base + off; };
  GArgs<T> a;
  a.n = n; a.m = m; a.f = S.f; a.lend = S.f + S.l; a.nsoc = S.nsoc();
  a.cg_max = (int)std::min<long>(std::max<long>(par.CG_MAX_ITS, 0), 100000);
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
  a.ax = (T*)at(o_ax); a.st = (T*)at(o_st); a.t = (T*)at(o_t);
  a.pq = (Pair<T>*)at(o_pq); a.u = (T*)at(o_u);
  a.p_res = (double*)at(o_pres); a.p_rho = (double*)at(o_prho);
  a.p_pu = (double*)at(o_ppu); a.p_q = (double*)at(o_pqn);
  a.sc = (GridScal*)at(o_sc);
  a.hist = hrows ? (double*)at(o_hist) : nullptr;
  a.hist_rows = hrows;
  std::memcpy(C.hbuf, P.buf.data(), P.buf.size());
  hipError_t e = hipMemcpyAsync(C.dbuf, C.hbuf, P.buf.size(), hipMemcpyHostToDevice, C.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
  if (e != hipSuccess) { err = std::string("upload: ") + hipGetErrorString(e); return false; }
  const double t1 = now_ms();
  hipLaunchKernelGGL(k_ax<T>, dim3(a.nb_a), dim3(GT), 0, C.stream, a);
  const long max_iters = std::max<long>(par.MAX_ITERS, 0);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (max_iters > 0) {
    e = hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      enqueue_iteration<T>(a, C.stream);
      e = hipStreamEndCapture(C.stream, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { err = std::string("graph: ") + hipGetErrorString(e); return false; }
  }
  GridScal* hsc = reinterpret_cast<GridScal*>(C.hbuf);
  long launched = 0;
  int chunk = 4;
  while (launched < max_iters) {
    const long k = std::min<long>(chunk, max_iters - launched);
    for (long i = 0; i < k && e == hipSuccess; ++i) e = hipGraphLaunch(exec, C.stream);
    launched += k;
    if (e == hipSuccess) e = hipMemcpyAsync(hsc, a.sc, sizeof(GridScal), hipMemcpyDeviceToHost, C.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
    if (e != hipSuccess || hsc->done) break;
    chunk = std::min(chunk * 2, 32);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hsc, a.sc, sizeof(GridScal), hipMemcpyDeviceToHost, C.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) { err = std::string("grid engine: ") + hipGetErrorString(e); return false; }
  const double t2 = now_ms();
  const GridScal scv = *hsc;
  std::vector<T> hx(n), hs(m), hy(m);
  e = hipMemcpy(hx.data(), a.x, n * sizeof(T), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hs.data(), a.s, m * sizeof(T), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hy.data(), a.y, m * sizeof(T), hipMemcpyDeviceToHost);
  if (e == hipSuccess && hist) {
    hist->assign((size_t)5 * hrows, 0.0);
    if (hrows) e = hipMemcpy(hist->data(), a.hist, sizeof(double) * 5 * hrows, hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) { err = std::string("download: ") + hipGetErrorString(e); return false; }
  for (int i = 0; i < n; ++i) S.x[i] = double(hx[i]);
  for (int j = 0; j < m; ++j) { S.s[j] = double(hs[j]); S.y[j] = double(hy[j]); }
  res.assign(1, DevResult{});
  DevResult& r = res[0];
  r.iters = scv.iters;
  r.status = scv.status;
  r.cg_iters = scv.cg_total;
  r.p_res = scv.pres; r.d_res = scv.dres; r.p_obj = scv.pobj; r.d_obj = scv.dobj; r.eta = scv.eta;
  upload_ms = t1 - t0;
  solve_ms = t2 - t1;
  return true;
}

}  // namespace

bool grid_preferred(const Setup& S, bool fp64) {
  const int lds = wg_lds_bytes(S.n, S.m, S.n + S.m, false, fp64 ? 8 : 4);
  return lds > ctx().max_lds || S.nnz() > 200000 || S.m > 16384;
}

bool solve_grid(Setup& S, const Params& par, bool fp64, int device, std::vector<DevResult>& res,
                double& upload_ms, double& solve_ms, std::string& err, std::vector<double>* hist) {
  return fp64 ? run_grid<double>(S, par, device, res, upload_ms, solve_ms, err, hist)
              : run_grid<float>(S, par, device, res, upload_ms, solve_ms, err, hist);
}

}  // namespace pdos_amd
