// batch_setup.hip -- the setup of a batch's instances on the device
// (common.h:69-215): the equilibration of A (NORMALIZE), lambda, the x1e3 of
// E and c, and the scaled values of A in the row and column forms the
// workgroup engine reads.  The host keeps the structural work (checks, the
// row pattern of A and its map to the CSC entries, the images' layout); the
// values are formed here with the host's arithmetic bit for bit: every sum
// is sequential in the host's order with separately rounded products and
// sums (contraction off), and the conversion to binary32 rounds to nearest
// as the host's does.
//
//   k_setup_group   one workgroup per distinct matrix: D (row sums of squares
//                   in row order, averaged over each second-order cone,
//                   clamped and inverted under a square root), the values
//                   times D, E from the column sums of their squares, the
//                   values times E and 1e3; then the image's value slots
//                   (row by row through the image's positions), D, E and,
//                   where the streaming variant may run, its copies.
//   k_setup_inst    one thread per instance: 1 + |b| and 1 + |c| of the
//                   caller's b and c, b times D, c times E0 (then 1e3),
//                   lambda = (1e-15 + |b~|) / (1e-15 + |c~|), the warm start
//                   scaled (y / D, x / E0 / 1e3, s D), written in place, and
//                   lambda and the stop-test scales into both copies of the
//                   instance's record.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "engine.hpp"

namespace pdos_amd {

namespace {

constexpr double kSetupRatio = 1e3;     // common.h:16
constexpr double kSetupRangeMax = 1e6;  // common.h:17
constexpr double kSetupRangeMin = 1e-6;  // common.h:18

__device__ __forceinline__ double setup_rsqrt(double v) {
#pragma clang fp contract(off)
  v = fabs(v);
  v = v > kSetupRangeMin ? v : kSetupRangeMin;
  v = v > kSetupRangeMax ? kSetupRangeMax : v;
  return 1.0 / sqrt(v);
}

template <typename T>
__global__ __launch_bounds__(256) void k_setup_group(const SetupGroup* __restrict__ groups) {
#pragma clang fp contract(off)
  const SetupGroup G = groups[blockIdx.x];
  const int tid = threadIdx.x, nt = blockDim.x;
  double* ax = G.ax;
  double* D = G.D;
  // D: row sums of squares in the order of the stable transpose (increasing
  // column, CSC order within a column), the mean over each cone's rows, then
  // 1 / sqrt of the clamped value
  if (G.normalize) {
    for (int i = tid; i < G.m; i += nt) {
      double a = 0.0;
      for (int q = G.rp[i]; q < G.rp[i + 1]; ++q) {
        const double v = ax[G.csr_src[q]];
        a += v * v;
      }
      D[i] = a;
    }
    __syncthreads();
    for (int c = tid; c < G.nsoc; c += nt) {
      const int r0 = G.soc[c], r1 = G.soc[c + 1];
      double acc = 0.0;
      for (int i = r0; i < r1; ++i) acc += D[i];
      const double mean = acc / (double)(long)(r1 - r0);
      for (int i = r0; i < r1; ++i) D[i] = mean;
    }
    __syncthreads();
    for (int i = tid; i < G.m; i += nt) D[i] = setup_rsqrt(D[i]);
    __syncthreads();
    // column j: values times D of their rows, then E0 from the column's sum
    // of squares in stored order, then the values times E0; the rows come
    // from the image's column positions and row indices where there is an
    // image (no CSC row array uploaded), else from the CSC row array
    for (int q = tid; q < G.n; q += nt) {
      int j, rb;
      if (G.cpos) {
        const int4 P = G.cpos[q];
        j = P.x;
        rb = P.y;
      } else {
        j = q;
        rb = G.cp[q];
      }
      const int p0 = G.cp[j], p1 = G.cp[j + 1];
      double a = 0.0;
      for (int p = p0; p < p1; ++p) {
        const int row = G.cpos ? G.cidx[rb + (p - p0)] : G.ci[p];
        const double v = ax[p] * D[row];
        ax[p] = v;
        a += v * v;
      }
      const double e = setup_rsqrt(a);
      G.E0[j] = e;
      for (int p = p0; p < p1; ++p) ax[p] = ax[p] * e;
    }
  } else {
    for (int i = tid; i < G.m; i += nt) D[i] = 1.0;
    for (int j = tid; j < G.n; j += nt) G.E0[j] = 1.0;
  }
  __syncthreads();
  for (int p = tid; p < G.nnz; p += nt) ax[p] = ax[p] * kSetupRatio;
  for (int j = tid; j < G.n; j += nt) {
    const double e = G.E0[j] * kSetupRatio;
    G.E[j] = e;
    if (G.E2) G.E2[j] = e;
  }
  if (G.D2)
    for (int i = tid; i < G.m; i += nt) G.D2[i] = D[i];
  __syncthreads();
  if (G.rval) {  // position q holds row (column) P.x's entries at [P.y, P.z) in their stored order
    T* rv = static_cast<T*>(G.rval);
    T* cv = static_cast<T*>(G.cval);
    for (int q = tid; q < G.m; q += nt) {
      const int4 P = G.rpos[q];
      const int r0 = G.rp[P.x];
      for (int e = 0; e < P.z - P.y; ++e) rv[P.y + e] = T(ax[G.csr_src[r0 + e]]);
    }
    for (int q = tid; q < G.n; q += nt) {
      const int4 P = G.cpos[q];
      const int c0 = G.cp[P.x];
      for (int e = 0; e < P.z - P.y; ++e) cv[P.y + e] = T(ax[c0 + e]);
    }
  }
  if (G.srx) {
    T* sr = static_cast<T*>(G.srx);
    T* sc = static_cast<T*>(G.scx);
    for (int q = tid; q < G.nnz; q += nt) sr[q] = T(ax[G.csr_src[q]]);
    for (int p = tid; p < G.nnz; p += nt) sc[p] = T(ax[p]);
  }
}

__global__ __launch_bounds__(256) void k_setup_inst(const SetupInst* __restrict__ insts, int count,
                                                    WgInst* __restrict__ rec, WgInst* __restrict__ rec_cls) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const SetupInst I = insts[t];
  const double* D = I.D;
  const double* E0 = I.E0;
  double nb = 0.0, nc = 0.0, nbs = 0.0, ncs = 0.0;
  for (int i = 0; i < I.m; ++i) {
    const double v = I.b[i];
    nb += v * v;
  }
  for (int j = 0; j < I.n; ++j) {
    const double v = I.c[j];
    nc += v * v;
  }
  for (int i = 0; i < I.m; ++i) {
    const double bs = I.b[i] * D[i];
    I.b[i] = bs;
    nbs += bs * bs;
  }
  for (int j = 0; j < I.n; ++j) {
    const double cs = I.c[j] * E0[j];
    ncs += cs * cs;
    I.c[j] = cs * kSetupRatio;
  }
  for (int i = 0; i < I.m; ++i) {
    I.y[i] = I.y[i] / D[i];
    I.s[i] = I.s[i] * D[i];
  }
  for (int j = 0; j < I.n; ++j) I.x[j] = (I.x[j] / E0[j]) / kSetupRatio;
  const double lam = (1e-15 + sqrt(nbs)) / (1e-15 + sqrt(ncs));
  const double ps = 1.0 + sqrt(nb), ds = 1.0 + sqrt(nc);
  I.out[0] = lam;
  I.out[1] = ps;
  I.out[2] = ds;
  rec[I.rec].lambda = lam;
  rec[I.rec].pscale = ps;
  rec[I.rec].dscale = ds;
  rec_cls[I.cls].lambda = lam;
  rec_cls[I.cls].pscale = ps;
  rec_cls[I.cls].dscale = ds;
}

}  // namespace

hipError_t batch_setup_launch(const SetupGroup* d_groups, int ngroups, const SetupInst* d_insts, int ninsts,
                              WgInst* rec, WgInst* rec_cls, bool fp64, hipStream_t st) {
  if (ngroups > 0) {
    if (fp64) hipLaunchKernelGGL(k_setup_group<double>, dim3(ngroups), dim3(256), 0, st, d_groups);
    else hipLaunchKernelGGL(k_setup_group<float>, dim3(ngroups), dim3(256), 0, st, d_groups);
  }
  if (ninsts > 0)
    hipLaunchKernelGGL(k_setup_inst, dim3((ninsts + 255) / 256), dim3(256), 0, st, d_insts, ninsts, rec, rec_cls);
  return hipGetLastError();
}

}  // namespace pdos_amd
