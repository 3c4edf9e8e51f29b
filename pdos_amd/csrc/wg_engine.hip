// wg_engine.hip -- the whole ADMM solve of one cone program in one workgroup.
//
// One workgroup (256 threads, 4 wave64) owns one instance for its whole solve:
// every iterate lives in LDS, matrices are read from global memory (L2/MALL
// resident for small instances), and the only synchronisation is the
// workgroup barrier.  A batch is one launch with one workgroup per instance
// (the hardware dispatcher keeps the CUs busy as instances retire); a single
// small problem is a launch of one workgroup.
//
// Per ADMM iteration (pdos.c:94-120) with k CG steps the workgroup runs
// k + 1 SpMVs with A and k + 2 with A' (the A x row sums of the final
// s~ = b - A x are reused by the next iteration's residual and CG start,
// which is exact: x does not change in between), and 4 + 3k barriers (4 + 2k
// when n <= 128: each wave then keeps all of x, q and p in registers and
// forms the CG step's dot products and updates itself).
//
// Reference map:
//   residual check      linAlg.h:158-201, pdos.c:95-107
//   CG on (I + A'A)      indirect/private.c:20-117 (p = beta p + q fused into
//                        the next A' sweep; the A sweep gathers beta p + q)
//   LDL' solve           direct/private.c:16-56, ldl.c:326-428 (level sets)
//   relax / cones / dual pdos.c:232-249, 296-303; cones.c:5-57
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>

#include "device.hpp"

namespace pdos_amd {

namespace {

constexpr int NT = kWgThreads;
constexpr int NW = NT / 64;

// Pointers into a group's image: in LDS (generic pointers derived from the
// kernel's LDS array, which the compiler resolves to LDS accesses) or in
// device memory, typed as global so that its loads are global loads, which
// the compiler may reorder around the LDS stores of the iterates (through
// generic pointers they would be flat loads, ordered against every LDS
// access).  The host pass only parses this code and keeps generic pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define PDOS_GLOBAL __attribute__((address_space(1)))
#else
#define PDOS_GLOBAL
#endif
template <bool IN_LDS>
struct ImgSpace {
  using Base = const unsigned char*;
  template <typename V>
  using Ptr = const V*;
};
template <>
struct ImgSpace<false> {
  using Base = const PDOS_GLOBAL unsigned char*;
  template <typename V>
  using Ptr = const PDOS_GLOBAL V*;
};
template <typename V, bool IN_LDS>
__device__ __forceinline__ typename ImgSpace<IN_LDS>::template Ptr<V> img_at(typename ImgSpace<IN_LDS>::Base b,
                                                                              int off) {
  return reinterpret_cast<typename ImgSpace<IN_LDS>::template Ptr<V>>(b + off);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sums NV doubles over the workgroup; every thread receives the totals.  One
// barrier; two alternating slots make back-to-back reductions race free (a
// slot is rewritten only after every thread has passed the next barrier).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red, int& slot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  double* buf = red + slot * (16 * 4);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) buf[w * 4 + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < NW; ++j) a += buf[j * 4 + k];
    v[k] = a;
  }
  slot ^= 1;
}

// SpMV over host-built row blocks {r0, r1, p0, p1} of whole rows (at most
// kWgChunk entries and kWgRowsPerBlock rows).  Per block every thread first
// loads its kWgChunk / NT entries and the row pointers it needs (coalesced,
// all in flight), forms the products val * gx(idx) into LDS, then the rows
// are summed from LDS and emit(row, sum) is called:
//   r1 > 0: rows shorter than kWgLongRow entries, one thread per row, in
//           entry order;
//   r1 < 0: rows [r0, -r1) of kWgLongRow entries or more, one wave per row
//           (lanes stride the row, then a wave reduction; emit on lane 0);
//   a single row longer than kWgChunk (p1 - p0 > kWgChunk) is summed by the
//   whole workgroup (emit on thread 0).
// Starts with a barrier: `prod` and the gathered vector may still be in use
// by the caller's previous phase.
template <typename T, typename Gather, typename Emit>
__device__ __forceinline__ void spmv_blocks(const int4* __restrict__ blk, int nblk, const int* __restrict__ ptr,
                                            const int* __restrict__ idx, const T* __restrict__ val, T* prod,
                                            double* red, int& slot, Gather gx, Emit emit) {
  constexpr int PT = kWgChunk / NT;
  constexpr int RT = kWgRowsPerBlock / NT;
  constexpr int LT = kWgChunk / kWgLongRow / NW;  // long rows per wave and block
  constexpr int QT = RT > LT ? RT : LT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int q0[QT], q1[QT], ci[PT];
  T cv[PT];
  // entries of block dd (skipped for a single row longer than kWgChunk)
  auto load_entries = [&](const int4& dd) {
    const int len = dd.w - dd.z;
    if (len > kWgChunk) return;
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * NT;
      if (e < len) { ci[u] = idx[dd.z + e]; cv[u] = val[dd.z + e]; }
    }
  };
  // the row pointers this thread (short rows) or wave (long rows) needs
  auto load_rows = [&](const int4& dd) {
    if (dd.w - dd.z > kWgChunk) return;
    if (dd.y >= 0) {
#pragma unroll
      for (int k = 0; k < RT; ++k) {
        const int r = dd.x + tid + k * NT;
        if (r < dd.y) { q0[k] = ptr[r] - dd.z; q1[k] = ptr[r + 1] - dd.z; }
      }
    } else {
#pragma unroll
      for (int k = 0; k < LT; ++k) {
        const int r = dd.x + w + k * NW;
        if (r < -dd.y) { q0[k] = ptr[r] - dd.z; q1[k] = ptr[r + 1] - dd.z; }
      }
    }
  };
  if (nblk == 0) return;
  int4 d = blk[0];
  load_entries(d);
  load_rows(d);
  for (int b = 0; b < nblk; ++b) {
    const bool more = b + 1 < nblk;
    const int len = d.w - d.z;
    if (len > kWgChunk) {
      double acc[1] = {0.0};
      for (int e = d.z + tid; e < d.w; e += NT) acc[0] += double(val[e] * gx(idx[e]));
      block_sum<1>(acc, red, slot);
      if (tid == 0) emit(d.x, T(acc[0]));
      if (more) {
        d = blk[b + 1];
        load_entries(d);
        load_rows(d);
      }
      continue;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int e = tid + u * NT;
      if (e < len) prod[e] = cv[u] * gx(ci[u]);
    }
    __syncthreads();
    // the next block's entries are in flight while this block's rows are summed
    int4 dn = d;
    if (more) {
      dn = blk[b + 1];
      load_entries(dn);
    }
    if (d.y >= 0) {
#pragma unroll
      for (int k = 0; k < RT; ++k) {
        const int r = d.x + tid + k * NT;
        if (r < d.y) {
          T acc = T(0);
#pragma unroll 4
          for (int e = q0[k]; e < q1[k]; ++e) acc += prod[e];
          emit(r, acc);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < LT; ++k) {
        const int r = d.x + w + k * NW;
        if (r < -d.y) {
          double acc = 0.0;
          for (int e = q0[k] + lane; e < q1[k]; e += 64) acc += double(prod[e]);
          acc = wave_sum(acc);
          if (lane == 0) emit(r, T(acc));
        }
      }
    }
    if (more) {
      load_rows(dn);
      d = dn;
    }
  }
}

// One level-scheduled sparse triangular sweep over LDS vector z.
// Levels with at most NW rows give each row a whole wave; wider levels give
// each row one thread (sequential in storage order).
template <bool BACK>
__device__ __forceinline__ void tri_sweep(int nlev, const int* __restrict__ lptr,
                                          const int* __restrict__ lrows, const int* __restrict__ ptr,
                                          const int* __restrict__ idx, const double* __restrict__ val,
                                          const double* __restrict__ dg, double* z) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int lv = 0; lv < nlev; ++lv) {
    const int a = lptr[lv], e = lptr[lv + 1];
    if (e - a <= NW) {
      if (a + w < e) {
        const int i = lrows[a + w];
        double acc = 0.0;
        for (int p = ptr[i] + lane; p < ptr[i + 1]; p += 64) acc += val[p] * z[idx[p]];
        acc = wave_sum(acc);
        if (lane == 0) z[i] = (BACK ? z[i] / dg[i] : z[i]) - acc;
      }
    } else {
      for (int rr = a + (int)threadIdx.x; rr < e; rr += NT) {
        const int i = lrows[rr];
        double v = BACK ? z[i] / dg[i] : z[i];
        for (int p = ptr[i]; p < ptr[i + 1]; ++p) v -= val[p] * z[idx[p]];
        z[i] = v;
      }
    }
    __syncthreads();
  }
}

// The streaming variant, for the instance of record `idx`: the iterates in
// LDS, the matrices (and factor) read from device memory.
template <typename T, bool DIRECT>
__device__ __forceinline__ void wg_stream_body(const WgInst* __restrict__ insts, int idx, unsigned char* smem) {
  // the record by value: its fields stay in registers across the barriers
  const WgInst I = insts[idx];
  const int n = I.n, m = I.m, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;

  // LDS carve-up (matches wg_lds_bytes)
  T* x = reinterpret_cast<T*>(smem);
  T *p = nullptr, *q = nullptr, *u = nullptr, *t = nullptr;
  T *s, *y, *st, *ax;
  double* bp = nullptr;
  unsigned char* tail;  // SpMV products, then the reduction slots
  if (!DIRECT) {
    p = x + n; q = p + n; u = q + n;
    s = u + n; y = s + m; st = y + m; t = st + m; ax = t + m;
    tail = smem + ((((size_t)(4 * n + 5 * m) * sizeof(T)) + 15) & ~(size_t)15);
  } else {
    s = x + n; y = s + m; st = y + m; ax = st + m;
    unsigned char* after = smem + ((((size_t)(n + 4 * m) * sizeof(T)) + 15) & ~(size_t)15);
    bp = reinterpret_cast<double*>(after);
    tail = after + (((size_t)I.N * 8 + 15) & ~(size_t)15);
  }
  T* prod = reinterpret_cast<T*>(tail);
  double* red = reinterpret_cast<double*>(tail + (((size_t)kWgChunk * sizeof(T) + 15) & ~(size_t)15));

  const T* __restrict__ rx = reinterpret_cast<const T*>(I.rx);
  const T* __restrict__ cx = reinterpret_cast<const T*>(I.cx);
  const int* __restrict__ rp = I.rp;
  const int* __restrict__ ri = I.ri;
  const int* __restrict__ cp = I.cp;
  const int* __restrict__ ci = I.ci;
  const double* __restrict__ bg = I.b;
  const double* __restrict__ cg = I.c;
  const double* __restrict__ Dg = I.D;
  const double* __restrict__ Eg = I.E;
  const int f = I.f, lend = I.f + I.l, nsoc = I.nsoc;
  const int4* __restrict__ rblk = reinterpret_cast<const int4*>(I.rblk);
  const int4* __restrict__ cblk = reinterpret_cast<const int4*>(I.cblk);
  const int nrb = I.nrb, ncb = I.ncb;
  int slot = 0;
  const T lam = T(I.lambda);
  const T relax_a = T(I.alpha);
  const T relax_b = T(1.0) - T(I.alpha);

  if (I.wsrc) {  // warm start in original units (common.h:148-161, 188): x/E, y/D, s*D
    const float* w = I.wsrc;
    for (int i = tid; i < n; i += NT) x[i] = T(double(w[i]) / Eg[i]);
    for (int j = tid; j < m; j += NT) {
      y[j] = T(double(w[n + j]) / Dg[j]);
      s[j] = T(double(w[n + m + j]) * Dg[j]);
    }
  } else {
    for (int i = tid; i < n; i += NT) x[i] = T(I.x0[i]);
    for (int j = tid; j < m; j += NT) { s[j] = T(I.s0[j]); y[j] = T(I.y0[j]); }
  }
  __syncthreads();
  spmv_blocks<T>(rblk, nrb, rp, ri, rx, prod, red, slot, [&](int c) { return x[c]; },
                 [&](int j, T v) { ax[j] = v; });
  __syncthreads();

  int it = 0, status = 1, cg_total = 0, fallback = 0;
  bool judge = I.warm_check != 0;
  double pres = -1.0, dres = -1.0, pobj = -1.0, dobj = -1.0, eta = -1.0;

  for (it = 0; it < I.max_iters; ++it) {
    // ---- stop check (linAlg.h:158-201); also the CG start argument t ----
    double acc[4] = {0.0, 0.0, 0.0, 0.0};  // |pres|^2, |dres|^2, c'x, b'y
    for (int j = tid; j < m; j += NT) {
      const T sj = s[j], bj = T(bg[j]), yj = y[j], axj = ax[j];
      const T r = ((sj - bj) + axj) / T(Dg[j]);
      acc[0] += double(r) * double(r);
      acc[3] += double(bj) * double(yj);
      if (!DIRECT) t[j] = ((sj + lam * yj) - bj) + axj;
    }
    spmv_blocks<T>(cblk, ncb, cp, ci, cx, prod, red, slot, [&](int c) { return y[c]; }, [&](int i, T v) {
      const T ci_ = T(cg[i]);
      const T r = (ci_ + v) / T(Eg[i]);
      acc[1] += double(r) * double(r);
      acc[2] += double(ci_) * double(x[i]);
    });
    block_sum<4>(acc, red, slot);
    pres = sqrt(acc[0]);
    dres = sqrt(acc[1]);
    pobj = acc[2];
    dobj = -acc[3];
    eta = fabs(pobj - dobj);
    if (judge) {
      // ADMM judges the warm start: keep it only if its scaled residuals are
      // no worse than the cold start's, ||b|| / pscale and ||c|| / dscale.
      judge = false;
      const double mw = fmax(fmax(pres / I.pscale, dres / I.dscale), eta / (1.0 + fabs(pobj) + fabs(dobj)));
      const double mc = fmax((I.pscale - 1.0) / I.pscale, (I.dscale - 1.0) / I.dscale);
      if (!(mw <= mc)) {
        fallback = 1;
        for (int i = tid; i < n; i += NT) x[i] = T(0);
        for (int j = tid; j < m; j += NT) { s[j] = T(0); y[j] = T(0); ax[j] = T(0); }
        __syncthreads();
        it = -1;
        continue;
      }
    }
    if (pres < I.eps * I.pscale && dres < I.eps * I.dscale &&
        eta < I.eps * (1.0 + fabs(pobj) + fabs(dobj))) {
      status = 0;
      break;
    }
    if (I.hist && tid == 0 && it % 10 == 0 && it / 10 < I.hist_rows) {
      double* h = I.hist + (size_t)(it / 10) * 5;
      h[0] = pres; h[1] = dres; h[2] = pobj; h[3] = dobj; h[4] = eta;
    }

    if (!DIRECT) {
      // ---- CG on (I + A'A) x = x - lam c + A'(b - s - lam y) ----
      double r1[1] = {0.0};
      spmv_blocks<T>(cblk, ncb, cp, ci, cx, prod, red, slot, [&](int c) { return t[c]; }, [&](int i, T v) {
        const T qi = (-lam) * T(cg[i]) - v;
        q[i] = qi;
        p[i] = T(0);
        r1[0] += double(qi) * double(qi);
      });
      block_sum<1>(r1, red, slot);
      double rho = r1[0];
      T beta = T(0);
      int k = 0;
      double qlast = 0.0;  // |q|^2 after the last step (indirect/private.c:48,85)
      if (rho > I.cg_tol_sq) {
        while (k < I.cg_max_its) {
          spmv_blocks<T>(rblk, nrb, rp, ri, rx, prod, red, slot, [&](int c) { return beta * p[c] + q[c]; },
                         [&](int j, T v) { t[j] = v; });
          __syncthreads();
          double r2[1] = {0.0};
          spmv_blocks<T>(cblk, ncb, cp, ci, cx, prod, red, slot, [&](int c) { return t[c]; }, [&](int i, T v) {
            const T pi = beta * p[i] + q[i];
            p[i] = pi;
            const T ui = pi + v;
            u[i] = ui;
            r2[0] += double(pi) * double(ui);
          });
          block_sum<1>(r2, red, slot);
          const T alpha = T(rho / r2[0]);
          double r3[1] = {0.0};
          for (int i = tid; i < n; i += NT) {
            x[i] += alpha * p[i];
            const T qi = q[i] - alpha * u[i];
            q[i] = qi;
            r3[0] += double(qi) * double(qi);
          }
          block_sum<1>(r3, red, slot);
          ++k;
          qlast = r3[0];
          if (r3[0] < I.cg_tol_sq) break;
          beta = T(r3[0] / rho);
          rho = r3[0];
        }
      }
      cg_total += k;
      if (I.cg_hist && tid == 0 && it < I.hist_rows * 10) {
        I.cg_hist[(size_t)it * 2] = double(k);
        I.cg_hist[(size_t)it * 2 + 1] = sqrt(qlast);
      }
    } else {
      // ---- LDL' solve of the KKT system, permuted into bp ----
      const int N = I.N;
      for (int kk = tid; kk < N; kk += NT) {
        const int src = I.perm[kk];
        double v;
        if (src < n) {
          v = double(lam) * cg[src] - double(x[src]);
        } else {
          const int j = src - n;
          v = bg[j] - (double(s[j]) + double(lam) * double(y[j]));
        }
        bp[kk] = v;
      }
      __syncthreads();
      tri_sweep<false>(I.nfl, I.fl_ptr, I.fl_rows, I.lrp, I.lri, I.lrx, nullptr, bp);
      tri_sweep<true>(I.nbl, I.bl_ptr, I.bl_rows, I.lcp, I.lci, I.lcx, I.ld, bp);
      for (int kk = tid; kk < N; kk += NT) {
        const int dst = I.perm[kk];
        const double v = bp[kk];
        if (dst < n) {
          x[dst] = T(v);
        } else {
          const int j = dst - n;
          st[j] = T(v + (double(s[j]) + double(lam) * double(y[j])));
        }
      }
      __syncthreads();
    }

    // ---- s~ = b - A x, relax, project, dual update (pdos.c:232-249,296-303) ----
    spmv_blocks<T>(rblk, nrb, rp, ri, rx, prod, red, slot, [&](int c) { return x[c]; }, [&](int j, T v) {
      ax[j] = v;
      T stj = DIRECT ? st[j] : T(bg[j]) - v;
      stj = relax_a * stj + relax_b * s[j];
      const T z = stj - lam * y[j];
      if (j < f) {
        s[j] = T(0);
        y[j] += (T(0) - stj) / lam;
      } else if (j < lend) {
        const T sj = z < T(0) ? T(0) : z;
        s[j] = sj;
        y[j] += (sj - stj) / lam;
      } else {
        st[j] = stj;
        s[j] = z;
      }
    });
    __syncthreads();
    for (int cn = wv; cn < nsoc; cn += NW) {
      const int r0 = I.soc_ptr[cn], r1 = I.soc_ptr[cn + 1];
      double a2 = 0.0;
      for (int j = r0 + 1 + lane; j < r1; j += 64) a2 += double(s[j]) * double(s[j]);
      a2 = wave_sum(a2);
      const T ns = T(sqrt(a2));
      const T t0 = s[r0];
      // 0: inside the cone, 1: inside the polar cone, 2: onto the boundary
      const int mode = ns <= t0 ? 0 : (ns <= -t0 ? 1 : 2);
      const T a = (ns + t0) / T(2);
      for (int j = r0 + lane; j < r1; j += 64) {
        const T zj = s[j];
        const T sj = mode == 0 ? zj : (mode == 1 ? T(0) : (j == r0 ? a : a * zj / ns));
        s[j] = sj;
        y[j] += (sj - st[j]) / lam;
      }
    }
    __syncthreads();
  }

  for (int i = tid; i < n; i += NT) I.x[i] = double(x[i]);
  for (int j = tid; j < m; j += NT) { I.s[j] = double(s[j]); I.y[j] = double(y[j]); }
  if (tid == 0) {
    DevResult* R = I.res;
    R->iters = it;
    R->status = status;
    R->cg_iters = cg_total;
    R->fallback = fallback;
    R->p_res = pres;
    R->d_res = dres;
    R->p_obj = pobj;
    R->d_obj = dobj;
    R->eta = eta;
  }
}

// ---------------------------------------------------------------------------
// Resident variant: for an instance whose matrices (and, direct plugin, the
// factor in level order) fit in LDS next to its iterates, the workgroup copies
// its group's image into LDS once and every SpMV, sweep and reduction of the
// solve touches LDS only -- no global memory round trip per phase or level.
//   SpMV and sweeps: rows in wave batches, 1 to 64 lanes per row so that no
//         lane sums more than 4 entries of a row of up to 256 (a row of one
//         lane in storage order, the reference's order);
//   direct: the permuted right-hand side is formed in the residual phase, the
//         sweeps run in level order (rows of a level stored contiguously),
//         and the un-permute is folded into the next A product (A's column
//         indices stored also as pivot positions), so an iteration is one
//         barrier per level plus three.

// Data-parallel-primitive (DPP) moves of a double between lanes, on its two
// 32-bit halves; masked-off rows read 0.  A move that writes every row (the
// quad permutes and mirrors) needs no prior value, so the destination is not
// zeroed first (two VALU instructions fewer per move).
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_move(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  unsigned lo, hi;
  if constexpr (ROWS == 0xf) {
    lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, false);
    hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  } else {
    lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xf, false);
    hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xf, false);
  }
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Sum over groups of G = 2^lg consecutive lanes; every lane of a group gets
// its group's sum.  Up to 16 lanes by DPP (quad permutes, then the half-row
// and row mirrors: ALU latency, no LDS), beyond that by lane shuffles.
__device__ __forceinline__ double group_sum(double v, int lg) {
  if (lg >= 1) v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  if (lg >= 2) v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  if (lg >= 3) v += dpp_move<0x141>(v);  // row_half_mirror
  if (lg >= 4) v += dpp_move<0x140>(v);  // row_mirror
  if (lg >= 5) v += __shfl_xor(v, 16, 64);
  if (lg >= 6) v += __shfl_xor(v, 32, 64);
  return v;
}

// Sum over the wave, by DPP: the 16-lane row sums, then row 1 += row 0 and
// row 3 += row 2 (row_bcast15), row 3 += row 1 (row_bcast31); lane 63 holds
// the total, which every lane reads.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  v += dpp_move<0x142, 0xA>(v);
  v += dpp_move<0x143, 0xC>(v);
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// block_sum with the DPP wave total; the waves' totals are then summed by
// the same DPP tree within each row of 16 lanes (lane j of every row reads
// wave j's total, zero past the last wave), so every lane ends with the
// block totals without a loop of dependent LDS reads and adds.
template <int NV>
__device__ __forceinline__ void block_total(double (&v)[NV], double* red, int& slot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_total(v[k]);
  double* buf = red + slot * (16 * 4);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) buf[w * 4 + k] = v[k];
  }
  __syncthreads();
  const int nw = blockDim.x >> 6, j = lane & 15;  // 4, 8 or 16 waves
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = j < nw ? buf[j * 4 + k] : 0.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = group_sum(v[k], 4);
  slot ^= 1;
}

// A lane's share of one row: entries e0, e0 + G, ... below e1.  The first
// four are loaded unconditionally (the entry arrays carry 256 zero entries of
// padding, so the reads stay inside them) and masked by selects, so the lane
// runs straight-line code with all loads of a kind in flight together; the
// rest (rows of more than 4 G entries) one by one.  Summed in entry order.
template <typename A, typename V, typename PI, typename PV, typename Gather>
__device__ __forceinline__ A lane_dot(PI idx, PV val, int e0, int e1, int G, Gather gx) {
  int ix[4];
  V vv[4], gv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    ix[j] = idx[e0 + j * G];
    vv[j] = val[e0 + j * G];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) gv[j] = gx(ix[j]);
  A acc = A(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) acc += e0 + j * G < e1 ? A(vv[j] * gv[j]) : A(0);
  for (int e = e0 + 4 * G; e < e1; e += G) acc += A(val[e] * gx(idx[e]));
  return acc;
}

// y = A x over rows held in LDS, by wave batches {first, count, lg}: lane l
// of a batch takes position first + (l >> lg), i.e. row pos.x, and every
// (1 << lg)-th of its entries [pos.y, pos.z); a row of one lane sums them
// in storage order.  The lane at the start of each group emits.
// A wave batch read from LDS, made wave-uniform (scalar registers), so the
// branches on its class are scalar branches, not exec-mask divergence.
template <typename PB>
__device__ __forceinline__ int4 uniform_batch(PB bat, int b) {
  const int4 B = bat[b];
  return int4{__builtin_amdgcn_readfirstlane(B.x), __builtin_amdgcn_readfirstlane(B.y),
              __builtin_amdgcn_readfirstlane(B.z), 0};
}

template <typename T, typename P4, typename PI, typename PV, typename Gather, typename Emit>
__device__ __forceinline__ void spmv_res(P4 pos, PI idx, PV val, P4 bat, int nbat, Gather gx, Emit emit) {
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  for (int b = w; b < nbat; b += nw) {
    const int4 B = uniform_batch(bat, b);
    const int k = lane >> B.z, sub = lane & ((1 << B.z) - 1);
    const bool on = k < B.y;
    const int4 P = pos[B.x + (on ? k : 0)];
    if (B.z == 0) {
      if (on) emit(P.x, lane_dot<T, T>(idx, val, P.y, P.z, 1, gx));
    } else {
      double acc = on ? lane_dot<double, T>(idx, val, P.y + sub, P.z, 1 << B.z, gx) : 0.0;
      acc = group_sum(acc, B.z);
      if (on && sub == 0) emit(P.x, T(acc));
    }
  }
}

// A wave's batch of a product (batch w; kSlots > 1 would hold w + nw, ...
// too: measured slower, at 256 threads as at 512) held in registers across
// the CG steps: per batch the lane's row, its (up to four) entries' indices
// and values and which of them exist, so a step's product is the gathers,
// the products and the group sum (the descriptor, position and entry loads
// are made once per solve).  Same arithmetic and order as spmv_res.
template <typename T>
struct RegRows {
  static constexpr int kSlots = 1;
  int lg[kSlots];  // lanes per row (log2), wave-uniform; -1: no batch in the slot
  int row[kSlots];
  bool on[kSlots], lead[kSlots];
  unsigned msk[kSlots];
  int ix[kSlots][4];
  T vv[kSlots][4];
};

// Loads the wave's batches of (bat, nbat) into R; returns false
// (wave-uniform) when they cannot be held this way: more than kSlots batches
// a wave, or a row of more than four entries a lane.
template <typename T, typename P4, typename PI, typename PV>
__device__ __forceinline__ bool reg_rows_load(P4 pos, PI idx, PV val, P4 bat, int nbat, RegRows<T>& R) {
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  bool fits = nbat <= RegRows<T>::kSlots * nw;
#pragma unroll
  for (int sl = 0; sl < RegRows<T>::kSlots; ++sl) {
    // the batches in reverse: their rows' lane groups grow with the batch
    // index, and the first waves are the first out of the previous phase
    const int b = nbat - 1 - (w + sl * nw);
    R.lg[sl] = -1;
    R.on[sl] = R.lead[sl] = false;
    R.msk[sl] = 0;
    R.row[sl] = 0;
    if (!fits || b < 0) continue;
    const int4 B = uniform_batch(bat, b);
    const int k = lane >> B.z, sub = lane & ((1 << B.z) - 1), G = 1 << B.z;
    R.lg[sl] = B.z;
    R.on[sl] = k < B.y;
    R.lead[sl] = R.on[sl] && sub == 0;
    const int4 P = pos[B.x + (R.on[sl] ? k : 0)];
    R.row[sl] = P.x;
    const int e0 = P.y + sub;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = e0 + j * G;
      R.ix[sl][j] = idx[e];
      R.vv[sl][j] = val[e];
      if (e < P.z) R.msk[sl] |= 1u << j;
    }
    fits = fits && __all(!R.on[sl] || e0 + 4 * G >= P.z);
  }
  return fits;
}

// The product over a wave's held batches: emit(row, value) on the lanes that
// spmv_res would emit from.  Every slot's gathers are issued before the
// first product.
template <typename T, typename Gather, typename Emit>
__device__ __forceinline__ void reg_rows_spmv(const RegRows<T>& R, Gather gx, Emit emit) {
  constexpr int S = RegRows<T>::kSlots;
  T gv[S][4];
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    if (R.lg[sl] < 0) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) gv[sl][j] = gx(R.ix[sl][j]);
  }
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    if (R.lg[sl] < 0) continue;
    if (R.lg[sl] == 0) {
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc += (R.msk[sl] >> j) & 1u ? T(R.vv[sl][j] * gv[sl][j]) : T(0);
      if (R.on[sl]) emit(R.row[sl], acc);
    } else {
      double acc = 0.0;
      if (R.on[sl]) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc += (R.msk[sl] >> j) & 1u ? double(R.vv[sl][j] * gv[sl][j]) : 0.0;
      }
      acc = group_sum(acc, R.lg[sl]);
      if (R.lead[sl]) emit(R.row[sl], T(acc));
    }
  }
}

// One triangular sweep in LDS, level by level (lb: level -> its batches;
// position q: pos[q] = {row, entries}; backward: divide by dq[q] first), one
// barrier per level.
template <bool BACK, typename PI, typename P4, typename PD>
__device__ __forceinline__ void res_sweep(int nlev, PI lb, P4 bat, P4 pos, PI idx, PD val, PD dq, double* z) {
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  auto zg = [&](int c) { return z[c]; };
  for (int lv = 0; lv < nlev; ++lv) {
    const int b1 = __builtin_amdgcn_readfirstlane(lb[lv + 1]);
    for (int b = __builtin_amdgcn_readfirstlane(lb[lv]) + w; b < b1; b += nw) {
      const int4 B = uniform_batch(bat, b);
      const int k = lane >> B.z, sub = lane & ((1 << B.z) - 1);
      const bool on = k < B.y;
      const int q = B.x + (on ? k : 0);
      const int4 P = pos[q];
      const double zi = z[P.x];
      const double d = BACK ? dq[q] : 1.0;
      double acc = on ? lane_dot<double, double>(idx, val, P.y + sub, P.z, 1 << B.z, zg) : 0.0;
      acc = group_sum(acc, B.z);
      if (on && sub == 0) z[P.x] = (BACK ? zi / d : zi) - acc;
    }
    __syncthreads();
  }
}

// The dense trailing chain T of the factor (t <= 64 rows): z_T <- M z_T with
// M = (L_T D_T L_T')^-1 formed on the host (tm, t x t row-major).  Wave 0:
// lane i stages z at row trow[i] in tz, then sums M(i, k) tz[k] over k in
// four interleaved partial sums (fixed order), every load issued before the
// first product.  One wave, so its staging writes precede its reads without
// a barrier; the caller's barrier publishes z_T.
template <typename PI, typename PD>
__device__ __forceinline__ void res_tblock(int t, PI trow, PD tm, double* z, double* tz) {
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const bool on = lane < t;
  const int r = on ? trow[lane] : 0;
  if (on) tz[lane] = z[r];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < t; k0 += 8) {
    double mv[8], zv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = k0 + u;
      mv[u] = on && k < t ? tm[lane * t + k] : 0.0;
      zv[u] = k < t ? tz[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u & 3] += mv[u] * zv[u];
  }
  if (on) z[r] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

// IMG_LDS = false: the same kernel with the group's image read from device
// memory (L2 / MALL resident) -- for instances whose matrices do not fit in
// LDS; only b, c, the iterates and the reduction slots are in LDS.
template <typename T, bool DIRECT, bool IMG_LDS, bool REGS = false>
__device__ __forceinline__ void wg_res_body(const WgInst* __restrict__ insts, int idx, unsigned char* smem) {
  // the record by value: its fields stay in registers across the barriers
  // (a reference would be re-read from memory after each one)
  const WgInst I = insts[idx];
  const ResLayout& L = I.lay;
  const int n = I.n, m = I.m, N = I.N, tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int nt = blockDim.x, nw = nt >> 6;  // 256, 512 or 1,024 threads
  if (IMG_LDS) {
    const uint4* src = reinterpret_cast<const uint4*>(I.img);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    for (int i = tid; i < L.img / 16; i += nt) dst[i] = src[i];
  }
  // static data: the image in LDS, or in device memory; per-instance data
  // after the image in LDS, or at the start of LDS
  typename ImgSpace<IMG_LDS>::Base sb;
  if constexpr (IMG_LDS) sb = smem;
  else sb = (typename ImgSpace<IMG_LDS>::Base)I.img;
  unsigned char* db = IMG_LDS ? smem : smem - L.img;
  const auto Dl = img_at<double, IMG_LDS>(sb, L.D);
  const auto El = img_at<double, IMG_LDS>(sb, L.E);
  const auto rpos = img_at<int4, IMG_LDS>(sb, L.rpos);
  const auto ridx = img_at<int, IMG_LDS>(sb, L.ridx);
  const auto rval = img_at<T, IMG_LDS>(sb, L.rval);
  const auto ridp = img_at<int, IMG_LDS>(sb, L.ridp);
  const auto cpos = img_at<int4, IMG_LDS>(sb, L.cpos);
  const auto cidx = img_at<int, IMG_LDS>(sb, L.cidx);
  const auto cval = img_at<T, IMG_LDS>(sb, L.cval);
  const auto rbat = img_at<int4, IMG_LDS>(sb, L.rbat);
  const auto cbat = img_at<int4, IMG_LDS>(sb, L.cbat);
  const int nrbat = L.nrbat, ncbat = L.ncbat;
  const auto socp = img_at<int, IMG_LDS>(sb, L.soc);
  const auto bigc = img_at<int, IMG_LDS>(sb, L.big);
  double* bl = reinterpret_cast<double*>(db + L.b);
  double* cl = reinterpret_cast<double*>(db + L.c);
  for (int j = tid; j < m; j += nt) bl[j] = I.b[j];
  for (int i = tid; i < n; i += nt) cl[i] = I.c[i];
  auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
  unsigned char* vp = db + L.vec;
  T* x = reinterpret_cast<T*>(vp);
  T* s = x + n;
  T* y = s + m;
  T* st = y + m;
  T* ax = st + m;
  vp += al((size_t)(n + 4 * m) * sizeof(T));
  T *p = nullptr, *q = nullptr, *u = nullptr, *t = nullptr;
  double *bp = nullptr, *tz = nullptr;
  if (DIRECT) {
    bp = reinterpret_cast<double*>(vp);
    vp += al((size_t)N * 8);
    tz = reinterpret_cast<double*>(vp);
    vp += 64 * 8;
  } else {
    p = reinterpret_cast<T*>(vp);
    q = p + n;
    u = q + n;
    t = u + n;
    vp += al((size_t)(3 * n + m) * sizeof(T));
  }
  double* red = reinterpret_cast<double*>(vp);
  const int f = I.f, lend = I.f + I.l, nsoc = I.nsoc;
  int slot = 0;
  const T lam = T(I.lambda);
  const T relax_a = T(I.alpha);
  const T relax_b = T(1.0) - T(I.alpha);
  __syncthreads();  // image, b and c in LDS

  if (I.wsrc) {  // warm start in original units (common.h:148-161, 188): x/E, y/D, s*D
    const float* w = I.wsrc;
    for (int i = tid; i < n; i += nt) x[i] = T(double(w[i]) / El[i]);
    for (int j = tid; j < m; j += nt) {
      y[j] = T(double(w[n + j]) / Dl[j]);
      s[j] = T(double(w[n + m + j]) * Dl[j]);
    }
  } else {
    for (int i = tid; i < n; i += nt) x[i] = T(I.x0[i]);
    for (int j = tid; j < m; j += nt) { s[j] = T(I.s0[j]); y[j] = T(I.y0[j]); }
  }
  __syncthreads();
  spmv_res<T>(rpos, ridx, rval, rbat, nrbat, [&](int c) { return x[c]; }, [&](int j, T v) { ax[j] = v; });
  __syncthreads();

  int it = 0, status = 1, cg_total = 0, fallback = 0;
  bool judge = I.warm_check != 0;
  double pres = -1.0, dres = -1.0, pobj = -1.0, dobj = -1.0, eta = -1.0;

  for (it = 0; it < I.max_iters; ++it) {
    // ---- stop check (linAlg.h:158-201); the CG start argument t or the
    // permuted KKT right-hand side (direct/private.c:16-31) ----
    double acc[4] = {0.0, 0.0, 0.0, 0.0};  // |pres|^2, |dres|^2, c'x, b'y
    for (int j = tid; j < m; j += nt) {
      const T sj = s[j], bj = T(bl[j]), yj = y[j], axj = ax[j];
      const T r = ((sj - bj) + axj) / T(Dl[j]);
      acc[0] += double(r) * double(r);
      acc[3] += double(bj) * double(yj);
      if (!DIRECT) t[j] = ((sj + lam * yj) - bj) + axj;
    }
    spmv_res<T>(cpos, cidx, cval, cbat, ncbat, [&](int c) { return y[c]; }, [&](int i, T v) {
      const T ci_ = T(cl[i]);
      const T r = (ci_ + v) / T(El[i]);
      acc[1] += double(r) * double(r);
      acc[2] += double(ci_) * double(x[i]);
    });
    if (DIRECT) {
      const auto perm = img_at<int, IMG_LDS>(sb, L.perm);
      for (int kk = tid; kk < N; kk += nt) {
        const int src = perm[kk];
        double v;
        if (src < n) {
          v = double(lam) * cl[src] - double(x[src]);
        } else {
          const int j = src - n;
          v = bl[j] - (double(s[j]) + double(lam) * double(y[j]));
        }
        bp[kk] = v;
      }
    }
    block_total<4>(acc, red, slot);
    pres = sqrt(acc[0]);
    dres = sqrt(acc[1]);
    pobj = acc[2];
    dobj = -acc[3];
    eta = fabs(pobj - dobj);
    if (judge) {
      judge = false;
      const double mw = fmax(fmax(pres / I.pscale, dres / I.dscale), eta / (1.0 + fabs(pobj) + fabs(dobj)));
      const double mc = fmax((I.pscale - 1.0) / I.pscale, (I.dscale - 1.0) / I.dscale);
      if (!(mw <= mc)) {
        fallback = 1;
        for (int i = tid; i < n; i += nt) x[i] = T(0);
        for (int j = tid; j < m; j += nt) { s[j] = T(0); y[j] = T(0); ax[j] = T(0); }
        __syncthreads();
        it = -1;
        continue;
      }
    }
    if (pres < I.eps * I.pscale && dres < I.eps * I.dscale &&
        eta < I.eps * (1.0 + fabs(pobj) + fabs(dobj))) {
      status = 0;
      break;
    }
    if (I.hist && tid == 0 && it % 10 == 0 && it / 10 < I.hist_rows) {
      double* h = I.hist + (size_t)(it / 10) * 5;
      h[0] = pres; h[1] = dres; h[2] = pobj; h[3] = dobj; h[4] = eta;
    }

    if (!DIRECT) {
      // ---- CG on (I + A'A) x = x - lam c + A'(b - s - lam y) ----
      double r1[1] = {0.0};
      spmv_res<T>(cpos, cidx, cval, cbat, ncbat, [&](int c) { return t[c]; }, [&](int i, T v) {
        const T qi = (-lam) * T(cl[i]) - v;
        q[i] = qi;
        p[i] = T(0);
        r1[0] += double(qi) * double(qi);
      });
      block_total<1>(r1, red, slot);
      double rho = r1[0];
      T beta = T(0);
      int k = 0;
      double qlast = 0.0;
      if (REGS && rho > I.cg_tol_sq && n <= 2 * 64) {
        // n of at most 128: every wave holds all of x, q and p in registers
        // (lane l rows l and l + 64) and forms the two dot products and the
        // vector updates itself, so a step has two barriers (after the A and
        // the A' product) and no cross-wave sum.  The waves compute the same
        // values in the same order, so they agree on every decision, and p
        // (the next A product's input) is written by all of them with equal
        // values: a wave reads back its own or an equal write, no barrier.
        T xr[2], qr[2], pr[2];
#pragma unroll
        for (int uu = 0; uu < 2; ++uu) {
          const int i = lane + 64 * uu;
          xr[uu] = i < n ? x[i] : T(0);
          qr[uu] = i < n ? q[i] : T(0);
          pr[uu] = qr[uu];  // beta p + q with p = 0
          if (i < n) p[i] = pr[uu];
        }
        // the wave's batch of each product in registers where it has at
        // most one and no row of more than four entries a lane
        RegRows<T> ra, rc;
        const bool hold_a = reg_rows_load<T>(rpos, ridx, rval, rbat, nrbat, ra);
        const bool hold_c = reg_rows_load<T>(cpos, cidx, cval, cbat, ncbat, rc);
        while (k < I.cg_max_its) {
          auto emit_t = [&](int j, T v) { t[j] = v; };
          auto emit_u = [&](int i, T v) { u[i] = v; };
          if (hold_a) reg_rows_spmv(ra, [&](int c) { return p[c]; }, emit_t);
          else spmv_res<T>(rpos, ridx, rval, rbat, nrbat, [&](int c) { return p[c]; }, emit_t);
          __syncthreads();
          if (hold_c) reg_rows_spmv(rc, [&](int c) { return t[c]; }, emit_u);
          else spmv_res<T>(cpos, cidx, cval, cbat, ncbat, [&](int c) { return t[c]; }, emit_u);
          __syncthreads();
          T ur[2];
          double r2 = 0.0;
#pragma unroll
          for (int uu = 0; uu < 2; ++uu) {
            const int i = lane + 64 * uu;
            ur[uu] = i < n ? pr[uu] + u[i] : T(0);
            r2 += double(pr[uu]) * double(ur[uu]);
          }
          r2 = wave_total(r2);
          const T alpha = T(rho / r2);
          double r3 = 0.0;
#pragma unroll
          for (int uu = 0; uu < 2; ++uu) {
            xr[uu] += alpha * pr[uu];
            qr[uu] = qr[uu] - alpha * ur[uu];
            r3 += double(qr[uu]) * double(qr[uu]);
          }
          r3 = wave_total(r3);
          ++k;
          qlast = r3;
          if (r3 < I.cg_tol_sq) break;
          beta = T(r3 / rho);
          rho = r3;
#pragma unroll
          for (int uu = 0; uu < 2; ++uu) {
            const int i = lane + 64 * uu;
            pr[uu] = beta * pr[uu] + qr[uu];
            if (i < n) p[i] = pr[uu];
          }
        }
        // x as every wave holds it: each wave writes all of it (equal
        // values), so its later reads see its own write or an equal one
#pragma unroll
        for (int uu = 0; uu < 2; ++uu) {
          const int i = lane + 64 * uu;
          if (i < n) x[i] = xr[uu];
        }
      } else if (rho > I.cg_tol_sq) {
        while (k < I.cg_max_its) {
          spmv_res<T>(rpos, ridx, rval, rbat, nrbat, [&](int c) { return beta * p[c] + q[c]; },
                      [&](int j, T v) { t[j] = v; });
          __syncthreads();
          double r2[1] = {0.0};
          spmv_res<T>(cpos, cidx, cval, cbat, ncbat, [&](int c) { return t[c]; }, [&](int i, T v) {
            const T pi = beta * p[i] + q[i];
            p[i] = pi;
            const T ui = pi + v;
            u[i] = ui;
            r2[0] += double(pi) * double(ui);
          });
          block_total<1>(r2, red, slot);
          const T alpha = T(rho / r2[0]);
          double r3[1] = {0.0};
          for (int i = tid; i < n; i += nt) {
            x[i] += alpha * p[i];
            const T qi = q[i] - alpha * u[i];
            q[i] = qi;
            r3[0] += double(qi) * double(qi);
          }
          block_total<1>(r3, red, slot);
          ++k;
          qlast = r3[0];
          if (r3[0] < I.cg_tol_sq) break;
          beta = T(r3[0] / rho);
          rho = r3[0];
        }
      }
      cg_total += k;
      if (I.cg_hist && tid == 0 && it < I.hist_rows * 10) {
        I.cg_hist[(size_t)it * 2] = double(k);
        I.cg_hist[(size_t)it * 2 + 1] = sqrt(qlast);
      }
      // ---- s~ = b - A x, relax, project, dual update (pdos.c:232-249,296-303) ----
      spmv_res<T>(rpos, ridx, rval, rbat, nrbat, [&](int c) { return x[c]; }, [&](int j, T v) {
        ax[j] = v;
        T stj = T(bl[j]) - v;
        stj = relax_a * stj + relax_b * s[j];
        const T z = stj - lam * y[j];
        if (j < f) {
          s[j] = T(0);
          y[j] += (T(0) - stj) / lam;
        } else if (j < lend) {
          const T sj = z < T(0) ? T(0) : z;
          s[j] = sj;
          y[j] += (sj - stj) / lam;
        } else {
          st[j] = stj;
          s[j] = z;
        }
      });
    } else {
      // ---- LDL' sweeps in level order, then x = z[0:n] and s~ = z[n:] +
      // (s + lam y) read through pivot positions inside the A product ----
      res_sweep<false>(L.nfl, img_at<int, IMG_LDS>(sb, L.flb), img_at<int4, IMG_LDS>(sb, L.fbat),
                       img_at<int4, IMG_LDS>(sb, L.fpos), img_at<int, IMG_LDS>(sb, L.fidx),
                       img_at<double, IMG_LDS>(sb, L.fval), img_at<double, IMG_LDS>(sb, L.fval), bp);
      if (L.t) {
        res_tblock(L.t, img_at<int, IMG_LDS>(sb, L.trow), img_at<double, IMG_LDS>(sb, L.tm), bp, tz);
        __syncthreads();
      }
      res_sweep<true>(L.nbl, img_at<int, IMG_LDS>(sb, L.blb), img_at<int4, IMG_LDS>(sb, L.bbat),
                      img_at<int4, IMG_LDS>(sb, L.bpos), img_at<int, IMG_LDS>(sb, L.bidx),
                      img_at<double, IMG_LDS>(sb, L.bval), img_at<double, IMG_LDS>(sb, L.bdq), bp);
      const auto pinv = img_at<int, IMG_LDS>(sb, L.pinv);
      for (int i = tid; i < n; i += nt) x[i] = T(bp[pinv[i]]);
      spmv_res<T>(rpos, ridp, rval, rbat, nrbat, [&](int c) { return T(bp[c]); }, [&](int j, T v) {
        ax[j] = v;
        T stj = T(bp[pinv[n + j]] + (double(s[j]) + double(lam) * double(y[j])));
        stj = relax_a * stj + relax_b * s[j];
        const T z = stj - lam * y[j];
        if (j < f) {
          s[j] = T(0);
          y[j] += (T(0) - stj) / lam;
        } else if (j < lend) {
          const T sj = z < T(0) ? T(0) : z;
          s[j] = sj;
          y[j] += (sj - stj) / lam;
        } else {
          st[j] = stj;
          s[j] = z;
        }
      });
    }
    __syncthreads();
    for (int cn = wv; cn < nsoc; cn += nw) {
      const int r0 = socp[cn], r1 = socp[cn + 1];
      if (r1 - r0 > kWgBigCone) continue;  // below, by the whole workgroup
      // a lane's entries j = r0 + lane + 64 u in chunks of four, each chunk's
      // loads issued together (stores to s and y would otherwise order every
      // load after the previous entry's stores)
      double a2 = 0.0;
      for (int j0 = r0 + 1 + lane; j0 < r1; j0 += 256) {
        T sv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sv[u] = j0 + 64 * u < r1 ? s[j0 + 64 * u] : T(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) a2 += double(sv[u]) * double(sv[u]);
      }
      a2 = wave_total(a2);
      const T ns = T(sqrt(a2));
      const T t0 = s[r0];
      const int mode = ns <= t0 ? 0 : (ns <= -t0 ? 1 : 2);
      const T a = (ns + t0) / T(2);
      for (int j0 = r0 + lane; j0 < r1; j0 += 256) {
        T zv[4], tv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + 64 * u;
          if (j < r1) { zv[u] = s[j]; tv[u] = st[j]; yv[u] = y[j]; }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + 64 * u;
          if (j < r1) {
            const T zj = zv[u];
            const T sj = mode == 0 ? zj : (mode == 1 ? T(0) : (j == r0 ? a : a * zj / ns));
            s[j] = sj;
            y[j] = yv[u] + (sj - tv[u]) / lam;
          }
        }
      }
    }
    // cones of more than kWgBigCone rows: each thread takes rows r0 + 1 + tid,
    // + nt, ... for the norm (one block reduction, which also orders every
    // read of the cone before the projection's writes), then rows r0 + tid,
    // + nt, ... for the projection.  The apex is read before the reduction.
    for (int cb = 0; cb < L.nbig; ++cb) {
      const int r0 = bigc[2 * cb], r1 = bigc[2 * cb + 1];
      const T t0 = s[r0];
      double a2[1] = {0.0};
      for (int j0 = r0 + 1 + tid; j0 < r1; j0 += 4 * nt) {
        T sv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) sv[u] = j0 + u * nt < r1 ? s[j0 + u * nt] : T(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) a2[0] += double(sv[u]) * double(sv[u]);
      }
      block_total<1>(a2, red, slot);
      const T ns = T(sqrt(a2[0]));
      const int mode = ns <= t0 ? 0 : (ns <= -t0 ? 1 : 2);
      const T a = (ns + t0) / T(2);
      for (int j0 = r0 + tid; j0 < r1; j0 += 4 * nt) {
        T zv[4], tv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + u * nt;
          if (j < r1) { zv[u] = s[j]; tv[u] = st[j]; yv[u] = y[j]; }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int j = j0 + u * nt;
          if (j < r1) {
            const T zj = zv[u];
            const T sj = mode == 0 ? zj : (mode == 1 ? T(0) : (j == r0 ? a : a * zj / ns));
            s[j] = sj;
            y[j] = yv[u] + (sj - tv[u]) / lam;
          }
        }
      }
    }
    __syncthreads();
  }

  for (int i = tid; i < n; i += nt) I.x[i] = double(x[i]);
  for (int j = tid; j < m; j += nt) { I.s[j] = double(s[j]); I.y[j] = double(y[j]); }
  if (tid == 0) {
    DevResult* R = I.res;
    R->iters = it;
    R->status = status;
    R->cg_iters = cg_total;
    R->fallback = fallback;
    R->p_res = pres;
    R->d_res = dres;
    R->p_obj = pobj;
    R->d_obj = dobj;
    R->eta = eta;
  }
}

// The workgroup engine: one instance per workgroup (queue == nullptr:
// instance blockIdx.x), or persistent workgroups taking instance after
// instance from the counter `queue` until it passes count (every workgroup
// reaches that exit).  Each record names its variant (WgInst::resident: 0
// streaming, 1 image in LDS, 2 image in device memory), so instances of every
// variant with the same LDS footprint share one launch.
// Registers are capped for four waves per SIMD (four workgroups per CU, the
// most the LDS classes use): the streaming body needs a few more and spills
// 2-3 values to scratch in binary32 (reloaded once per CG step), about a dozen
// in binary64.
template <typename T, bool DIRECT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_wg(const WgInst* __restrict__ insts,
                                                                                int count, int* queue) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ int wg_next;
  int idx = blockIdx.x;
  for (;;) {
    if (queue) {
      __syncthreads();  // the previous instance is done with LDS and wg_next
      if (threadIdx.x == 0) wg_next = atomicAdd(queue, 1);
      __syncthreads();
      idx = wg_next;
    }
    if (idx >= count) break;
    const int v = __builtin_amdgcn_readfirstlane(insts[idx].resident);
    if (v == 1) wg_res_body<T, DIRECT, true>(insts, idx, smem);
    else if (v == 2) wg_res_body<T, DIRECT, false>(insts, idx, smem);
    else wg_stream_body<T, DIRECT>(insts, idx, smem);
    if (!queue) break;
  }
}

// The image variants with 512 or 1,024 threads per workgroup (8 or 16
// waves): where the LDS allows few workgroups per CU, more waves per
// workgroup keep more of the CU's SIMDs busy on the same instance (rows,
// batches and levels dealt over more waves).  Records of the streaming
// variant never come here (its row blocks are sized for 256 threads).
template <typename T, bool DIRECT, bool REGS>
__device__ __forceinline__ void wg_image_loop(const WgInst* __restrict__ insts, int count, int* queue,
                                              unsigned char* smem, int* wg_next) {
  int idx = blockIdx.x;
  for (;;) {
    if (queue) {
      __syncthreads();  // the previous instance is done with LDS and wg_next
      if (threadIdx.x == 0) *wg_next = atomicAdd(queue, 1);
      __syncthreads();
      idx = *wg_next;
    }
    if (idx >= count) break;
    const int v = __builtin_amdgcn_readfirstlane(insts[idx].resident);
    if (v == 1) wg_res_body<T, DIRECT, true, REGS>(insts, idx, smem);
    else wg_res_body<T, DIRECT, false, REGS>(insts, idx, smem);
    if (!queue) break;
  }
}

template <typename T, bool DIRECT>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4))) void k_wg_wide(
    const WgInst* __restrict__ insts, int count, int* queue) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ int wg_next;
  wg_image_loop<T, DIRECT, false>(insts, count, queue, smem, &wg_next);
}

// The image variants at up to 512 threads when every workgroup has a CU of
// its own (a single solve, or no more instances than CUs): at most two waves
// per SIMD, so the body may hold 256 registers a lane, and the CG steps of
// an instance with n <= 128 keep x, q, p and the wave's matrix rows in
// registers.
template <typename T, bool DIRECT>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void k_wg_solo(
    const WgInst* __restrict__ insts, int count, int* queue) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ int wg_next;
  wg_image_loop<T, DIRECT, true>(insts, count, queue, smem, &wg_next);
}

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0) p[0] = 0;
}

// A kernel's dynamic LDS limit, raised only (per process and kernel).
hipError_t lds_limit(const void* fn, int lds_bytes, std::atomic<int>& set) {
  if (lds_bytes <= set.load()) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
  if (e != hipSuccess) return e;
  int cur = set.load();
  while (lds_bytes > cur && !set.compare_exchange_weak(cur, lds_bytes)) {
  }
  return hipSuccess;
}

template <typename T, bool DIRECT>
hipError_t launch_t(const WgInst* d_insts, int count, int lds_bytes, int threads, int* queue, int grid, bool solo,
                    hipStream_t st) {
  static std::atomic<int> lds_set{-1}, lds_set_wide{-1}, lds_set_solo{-1};
  const bool wide = threads > NT;
  solo = solo && threads <= 512;
  auto fn = solo ? k_wg_solo<T, DIRECT> : (wide ? k_wg_wide<T, DIRECT> : k_wg<T, DIRECT>);
  hipError_t e = lds_limit(reinterpret_cast<const void*>(fn), lds_bytes,
                           solo ? lds_set_solo : (wide ? lds_set_wide : lds_set));
  if (e != hipSuccess) return e;
  if (queue) {  // a kernel, not a memset node, so that the launch may be captured in a graph
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, queue);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(fn, dim3(queue ? std::max(1, std::min(grid, count)) : count),
                     dim3(wide || solo ? threads : NT),
                     lds_bytes, st, d_insts, count, queue);
  return hipGetLastError();
}

// Plain device copy, 16 bytes per thread and step (the tail bytewise).
__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16,
                       const unsigned char* __restrict__ src_b, unsigned char* __restrict__ dst_b, size_t tail) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tail; i += stride) dst_b[i] = src_b[i];
}

}  // namespace

// dst = src (device memory, `bytes`), by a kernel of this library: the
// reference kernel for scaling memory-traffic counters.
hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const size_t n16 = bytes / 16, tail = bytes % 16;
  const unsigned char* sb = static_cast<const unsigned char*>(src) + n16 * 16;
  unsigned char* db = static_cast<unsigned char*>(dst) + n16 * 16;
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, st, static_cast<const uint4*>(src), static_cast<uint4*>(dst),
                     n16, sb, db, tail);
  return hipGetLastError();
}

// Launches the workgroup engine over `count` instances (one workgroup each,
// each in the variant its record names; threads 256, or 512 / 1,024 for a
// class of image-variant records only); queue (device int) != nullptr:
// `grid` persistent workgroups drain the records in order through that
// counter (zeroed here first).
hipError_t wg_launch(const WgInst* d_insts, int count, int lds_bytes, int threads, bool fp64, bool direct, int* queue,
                     int grid, bool solo, hipStream_t st) {
  if (threads != 256 && threads != 512 && threads != 1024) return hipErrorInvalidValue;
  if (fp64) return direct ? launch_t<double, true>(d_insts, count, lds_bytes, threads, queue, grid, solo, st)
                          : launch_t<double, false>(d_insts, count, lds_bytes, threads, queue, grid, solo, st);
  return direct ? launch_t<float, true>(d_insts, count, lds_bytes, threads, queue, grid, solo, st)
                : launch_t<float, false>(d_insts, count, lds_bytes, threads, queue, grid, solo, st);
}

}  // namespace pdos_amd
