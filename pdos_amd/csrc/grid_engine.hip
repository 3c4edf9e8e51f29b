// grid_engine.hip -- one large cone program on the whole GPU.
//
// Each phase of the ADMM iteration is one grid-stride kernel (2048 workgroups
// of 256 threads).  Reductions never leave the device: every workgroup writes
// its partial sums, and the last workgroup to finish (device-scope atomic
// ticket) sums the partials in a fixed order and updates the scalar record
// (rho, alpha, beta, flags, residuals).  Kernels read that record at entry and
// return at once when the solve has stopped or CG has converged, so one ADMM
// iteration is a fixed sequence of launches, captured once into a hipGraph
// and replayed; the host only polls the stop flag between batches of
// iterations.  The stop iteration is exact.
//
// Launches per iteration: resid (A' SpMV + vector pass), cg_init (A' SpMV),
// CG_MAX_ITS x {A SpMV, A' SpMV, vector pass}, fin (A SpMV + relax +
// nonneg cones), cone (SOCs + dual update).  SpMVs executed: k+1 with A and
// k+2 with A' for k CG steps (indirect/private.c:31-117 needs k+3 and k+2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "engine.hpp"

namespace pdos_amd {

namespace {

constexpr int GT = 256;
constexpr int GW = GT / 64;
constexpr int kGridBlocks = 2048;

struct GridScal {
  double rho, alpha, beta;
  double pres, dres, pobj, dobj, eta;
  int admm_done, status, iters, cg_active, cg_k, cg_total;
  unsigned int ticket;
  int pad;
};

template <typename T>
struct GArgs {
  int n, m, f, lend, nsoc, ga, gat, cg_max;
  double eps, relax, tol_sq, lambda, pscale, dscale;
  const int* rp; const int* ri; const T* rx;
  const int* cp; const int* ci; const T* cx;
  const double* b; const double* c; const double* D; const double* E;
  const int* soc_ptr;
  T* x; T* s; T* y; T* ax; T* st; T* t; T* p; T* q; T* u;
  double* part;
  GridScal* sc;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T group_sum(T v, int G) {
  for (int o = G >> 1; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Grid-stride row gather with G lanes per row (block-uniform loop bound).
template <typename T, typename Gather, typename Emit>
__device__ __forceinline__ void grid_rows(int nrows, const int* __restrict__ ptr,
                                          const int* __restrict__ idx, const T* __restrict__ val,
                                          int G, Gather gx, Emit emit) {
  const int g = threadIdx.x & (G - 1);
  const int grp = threadIdx.x / G, ngrp = GT / G;
  for (int base = blockIdx.x * ngrp; base < nrows; base += gridDim.x * ngrp) {
    const int r = base + grp;
    T acc = T(0);
    if (r < nrows) {
      const int pe = ptr[r + 1];
      for (int p = ptr[r] + g; p < pe; p += G) acc += val[p] * gx(idx[p]);
    }
    acc = group_sum(acc, G);
    if (r < nrows && g == 0) emit(r, acc);
  }
}

// Writes this block's partial sums; returns true in the last block to
// arrive, with the fixed-order grid totals in tot (every thread of it).
template <int NV>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], double* part, unsigned* ticket,
                                            double (&tot)[NV]) {
  __shared__ double sred[GW][4];
  __shared__ int s_last;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sred[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = 0.0;
#pragma unroll
      for (int j = 0; j < GW; ++j) a += sred[j][k];
      part[blockIdx.x * 4 + k] = a;
    }
    __threadfence();
    const unsigned prev = atomicAdd(ticket, 1u);
    s_last = prev == gridDim.x - 1;
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  double a[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) a[k] = 0.0;
  for (int bl = threadIdx.x; bl < (int)gridDim.x; bl += GT) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      a[k] += __hip_atomic_load(&part[bl * 4 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) a[k] = wave_sum(a[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sred[w][k] = a[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < GW; ++j) s += sred[j][k];
    tot[k] = s;
  }
  if (threadIdx.x == 0) *ticket = 0u;
  return true;
}

template <typename T>
__global__ __launch_bounds__(GT) void k_ax(GArgs<T> a) {
  grid_rows<T>(a.m, a.rp, a.ri, a.rx, a.ga, [&](int c) { return a.x[c]; },
               [&](int j, T v) { a.ax[j] = v; });
}

// Stop check (linAlg.h:158-201, pdos.c:95-107) + CG start argument t.
template <typename T>
__global__ __launch_bounds__(GT) void k_resid(GArgs<T> a) {
  GridScal* sc = a.sc;
  if (sc->admm_done) return;
  const T lam = T(a.lambda);
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = blockIdx.x * GT + threadIdx.x; j < a.m; j += gridDim.x * GT) {
    const T sj = a.s[j], bj = T(a.b[j]), yj = a.y[j], axj = a.ax[j];
    const T r = ((sj - bj) + axj) / T(a.D[j]);
    v[0] += double(r) * double(r);
    v[3] += double(bj) * double(yj);
    a.t[j] = ((sj + lam * yj) - bj) + axj;
  }
  grid_rows<T>(a.n, a.cp, a.ci, a.cx, a.gat, [&](int c) { return a.y[c]; }, [&](int i, T w) {
    const T ci = T(a.c[i]);
    const T r = (ci + w) / T(a.E[i]);
    v[1] += double(r) * double(r);
    v[2] += double(ci) * double(a.x[i]);
  });
  double tot[4];
  if (grid_reduce<4>(v, a.part, &sc->ticket, tot) && threadIdx.x == 0) {
    const double pres = sqrt(tot[0]), dres = sqrt(tot[1]), pobj = tot[2], dobj = -tot[3];
    const double eta = fabs(pobj - dobj);
    sc->pres = pres; sc->dres = dres; sc->pobj = pobj; sc->dobj = dobj; sc->eta = eta;
    if (pres < a.eps * a.pscale && dres < a.eps * a.dscale &&
        eta < a.eps * (1.0 + fabs(pobj) + fabs(dobj))) {
      sc->admm_done = 1;
      sc->status = 0;
    } else {
      sc->iters += 1;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(GT) void k_cginit(GArgs<T> a) {
  GridScal* sc = a.sc;
  if (sc->admm_done) return;
  const T lam = T(a.lambda);
  double v[1] = {0.0};
  grid_rows<T>(a.n, a.cp, a.ci, a.cx, a.gat, [&](int c) { return a.t[c]; }, [&](int i, T w) {
    const T qi = (-lam) * T(a.c[i]) - w;
    a.q[i] = qi;
    a.p[i] = T(0);
    v[0] += double(qi) * double(qi);
  });
  double tot[1];
  if (grid_reduce<1>(v, a.part, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->rho = tot[0];
    sc->beta = 0.0;
    sc->cg_k = 0;
    sc->cg_active = (tot[0] > a.tol_sq && a.cg_max > 0) ? 1 : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(GT) void k_cg_a(GArgs<T> a) {
  GridScal* sc = a.sc;
  if (sc->admm_done || !sc->cg_active) return;
  const T beta = T(sc->beta);
  grid_rows<T>(a.m, a.rp, a.ri, a.rx, a.ga, [&](int c) { return beta * a.p[c] + a.q[c]; },
               [&](int j, T w) { a.t[j] = w; });
}

template <typename T>
__global__ __launch_bounds__(GT) void k_cg_at(GArgs<T> a) {
  GridScal* sc = a.sc;
  if (sc->admm_done || !sc->cg_active) return;
  const T beta = T(sc->beta);
  double v[1] = {0.0};
  grid_rows<T>(a.n, a.cp, a.ci, a.cx, a.gat, [&](int c) { return a.t[c]; }, [&](int i, T w) {
    const T pi = beta * a.p[i] + a.q[i];
    a.p[i] = pi;
    const T ui = pi + w;
    a.u[i] = ui;
    v[0] += double(pi) * double(ui);
  });
  double tot[1];
  if (grid_reduce<1>(v, a.part, &sc->ticket, tot) && threadIdx.x == 0) sc->alpha = sc->rho / tot[0];
}

template <typename T>
__global__ __launch_bounds__(GT) void k_cg_upd(GArgs<T> a) {
  GridScal* sc = a.sc;
  if (sc->admm_done || !sc->cg_active) return;
  const T alpha = T(sc->alpha);
  double v[1] = {0.0};
  for (int i = blockIdx.x * GT + threadIdx.x; i < a.n; i += gridDim.x * GT) {
    a.x[i] += alpha * a.p[i];
    const T qi = a.q[i] - alpha * a.u[i];
    a.q[i] = qi;
    v[0] += double(qi) * double(qi);
  }
  double tot[1];
  if (grid_reduce<1>(v, a.part, &sc->ticket, tot) && threadIdx.x == 0) {
    const double rn = tot[0];
    const int k = sc->cg_k + 1;
    sc->cg_k = k;
    sc->cg_total += 1;
    if (rn < a.tol_sq || k >= a.cg_max) {
      sc->cg_active = 0;
    } else {
      sc->beta = rn / sc->rho;
      sc->rho = rn;
    }
  }
}

// s~ = b - A x, relax, zero / nonneg cones and their dual update; SOC rows
// keep s~ and z = s~ - lambda y for k_cone (pdos.c:232-249, 296-303).
template <typename T>
__global__ __launch_bounds__(GT) void k_fin(GArgs<T> a) {
  if (a.sc->admm_done) return;
  const T lam = T(a.lambda), ra = T(a.relax), rb = T(1.0) - T(a.relax);
  grid_rows<T>(a.m, a.rp, a.ri, a.rx, a.ga, [&](int c) { return a.x[c]; }, [&](int j, T w) {
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
  });
}

// One wave per second-order cone (cones.c:26-47) + dual update.
template <typename T>
__global__ __launch_bounds__(GT) void k_cone(GArgs<T> a) {
  if (a.sc->admm_done) return;
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
void enqueue_iteration(const GArgs<T>& a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(k_resid<T>, dim3(grid), dim3(GT), 0, st, a);
  hipLaunchKernelGGL(k_cginit<T>, dim3(grid), dim3(GT), 0, st, a);
  for (int k = 0; k < a.cg_max; ++k) {
    hipLaunchKernelGGL(k_cg_a<T>, dim3(grid), dim3(GT), 0, st, a);
    hipLaunchKernelGGL(k_cg_at<T>, dim3(grid), dim3(GT), 0, st, a);
    hipLaunchKernelGGL(k_cg_upd<T>, dim3(grid), dim3(GT), 0, st, a);
  }
  hipLaunchKernelGGL(k_fin<T>, dim3(grid), dim3(GT), 0, st, a);
  hipLaunchKernelGGL(k_cone<T>, dim3(grid), dim3(GT), 0, st, a);
}

template <typename T>
size_t put_as(Packer& P, const std::vector<double>& v) {
  std::vector<T> w(v.begin(), v.end());
  return P.put(w.data(), w.size() * sizeof(T));
}

template <typename T>
bool run_grid(Setup& S, const Params& par, int device, std::vector<DevResult>& res,
              double& upload_ms, double& solve_ms, std::string& err) {
  Ctx& C = ctx();
  if (!C.ensure(device, err)) return false;
  const double t0 = now_ms();
  const int n = S.n, m = S.m;
  Packer P;
  const size_t o_rp = P.put(S.rp.data(), S.rp.size() * 4), o_ri = P.put(S.ri.data(), S.ri.size() * 4);
  const size_t o_rx = put_as<T>(P, S.rx);
  const size_t o_cp = P.put(S.cp.data(), S.cp.size() * 4), o_ci = P.put(S.ci.data(), S.ci.size() * 4);
  const size_t o_cx = put_as<T>(P, S.cx);
  const size_t o_b = P.put(S.b.data(), m * 8), o_c = P.put(S.c.data(), n * 8);
  const size_t o_D = P.put(S.D.data(), m * 8), o_E = P.put(S.E.data(), n * 8);
  const size_t o_soc = P.put(S.soc_ptr.data(), S.soc_ptr.size() * 4);
  const size_t o_x = put_as<T>(P, S.x), o_s = put_as<T>(P, S.s), o_y = put_as<T>(P, S.y);
  const size_t o_ax = P.zeros(m * sizeof(T)), o_st = P.zeros(m * sizeof(T)), o_t = P.zeros(m * sizeof(T));
  const size_t o_p = P.zeros(n * sizeof(T)), o_q = P.zeros(n * sizeof(T)), o_u = P.zeros(n * sizeof(T));
  const size_t o_part = P.zeros(kGridBlocks * 4 * sizeof(double));
  GridScal sc0;
  std::memset(&sc0, 0, sizeof(sc0));
  sc0.status = 1;
  sc0.pres = sc0.dres = sc0.pobj = sc0.dobj = sc0.eta = -1.0;
  const size_t o_sc = P.put(&sc0, sizeof(sc0));
  if (!C.reserve(P.buf.size(), std::max(P.buf.size(), (size_t)(n + 2 * m) * 8 + 4096), err)) return false;
  unsigned char* base = static_cast<unsigned char*>(C.dbuf);
  auto at = [&](size_t off) { return base + off; };
  GArgs<T> a;
  a.n = n; a.m = m; a.f = S.f; a.lend = S.f + S.l; a.nsoc = S.nsoc();
  a.ga = lanes_for(S.nnz(), m);
  a.gat = lanes_for(S.nnz(), n);
  a.cg_max = (int)std::min<long>(std::max<long>(par.CG_MAX_ITS, 0), 100000);
  a.eps = par.EPS_ABS; a.relax = par.ALPHA; a.tol_sq = par.CG_TOL * par.CG_TOL;
  a.lambda = S.lambda; a.pscale = S.pscale; a.dscale = S.dscale;
  a.rp = (const int*)at(o_rp); a.ri = (const int*)at(o_ri); a.rx = (const T*)at(o_rx);
  a.cp = (const int*)at(o_cp); a.ci = (const int*)at(o_ci); a.cx = (const T*)at(o_cx);
  a.b = (const double*)at(o_b); a.c = (const double*)at(o_c);
  a.D = (const double*)at(o_D); a.E = (const double*)at(o_E);
  a.soc_ptr = (const int*)at(o_soc);
  a.x = (T*)at(o_x); a.s = (T*)at(o_s); a.y = (T*)at(o_y);
  a.ax = (T*)at(o_ax); a.st = (T*)at(o_st); a.t = (T*)at(o_t);
  a.p = (T*)at(o_p); a.q = (T*)at(o_q); a.u = (T*)at(o_u);
  a.part = (double*)at(o_part);
  a.sc = (GridScal*)at(o_sc);
  std::memcpy(C.hbuf, P.buf.data(), P.buf.size());
  hipError_t e = hipMemcpyAsync(C.dbuf, C.hbuf, P.buf.size(), hipMemcpyHostToDevice, C.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
  if (e != hipSuccess) { err = std::string("upload: ") + hipGetErrorString(e); return false; }
  const double t1 = now_ms();
  const int grid = kGridBlocks;
  hipLaunchKernelGGL(k_ax<T>, dim3(grid), dim3(GT), 0, C.stream, a);
  const long max_iters = std::max<long>(par.MAX_ITERS, 0);
  // one iteration as a graph, replayed; host polls the stop flag per chunk
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (max_iters > 0) {
    e = hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      enqueue_iteration<T>(a, grid, C.stream);
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
    if (e != hipSuccess || hsc->admm_done) break;
    chunk = std::min(chunk * 2, 64);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hsc, a.sc, sizeof(GridScal), hipMemcpyDeviceToHost, C.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e != hipSuccess) { err = std::string("grid engine: ") + hipGetErrorString(e); return false; }
  const double t2 = now_ms();
  const GridScal scv = *hsc;
  // iterates back to binary64
  std::vector<T> hx(n), hs(m), hy(m);
  e = hipMemcpy(hx.data(), a.x, n * sizeof(T), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hs.data(), a.s, m * sizeof(T), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(hy.data(), a.y, m * sizeof(T), hipMemcpyDeviceToHost);
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
                double& upload_ms, double& solve_ms, std::string& err) {
  return fp64 ? run_grid<double>(S, par, device, res, upload_ms, solve_ms, err)
              : run_grid<float>(S, par, device, res, upload_ms, solve_ms, err);
}

}  // namespace pdos_amd
