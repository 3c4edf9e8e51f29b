// dense_gemm.hip -- the dense shared-A family's two products per ADMM
// iteration on BF16 MFMA, hand-written for this shape (linAlg.h:70-115 for a
// batch of right-hand sides):
//
//   A' product  U = A' Y and V = A' W   (n x cols each, K = m: tall-skinny)
//   A product   Z = A X                 (m x cols, K = n)
//
// Both run as three bf16 products with fp32 accumulation on the split
// operands a = a_hi + a_lo, x = x_hi + x_lo (bf16 each): a x ~ a_hi x_hi +
// a_hi x_lo + a_lo x_hi, the dropped a_lo x_lo being ~2^-16 of the terms, so
// the products carry ~fp32 accuracy at bf16 MFMA rate
// (v_mfma_f32_16x16x32_bf16).
//
// Operands: A's split copies are stored once at setup in the layout the MFMA
// A operand wants -- a lane's 8 consecutive K values contiguous -- so they are
// loaded from memory straight into registers (16 bytes a lane): A' rows
// (n x m) for the A' product, A rows (m x n) for the A product, both padded
// with zeros to multiples of 64.  The right-hand sides are the engine's
// feature-major binary32 columns (K rows x ldb instances, instances
// contiguous); each thread loads 8 consecutive rows of one column (every
// load coalesced across the wave's columns), splits them into bf16 hi / lo
// and writes each as one 16-byte LDS store, so the MFMA B operand is read
// from LDS as one 16-byte load a lane as well.
//
// Tiles: a workgroup of 4 waves (4 x 1) computes a 128 x 64 output tile (per
// right-hand side), each wave 32 rows x 64 columns, so a wave's A fragments
// serve four column subtiles (a 64 x 64 tile of 2 x 2 waves loaded each A
// fragment twice and moved twice the bytes per product); the A' product
// splits K (= m) over blockIdx.z when the batch is narrow, each split
// writing its own plane of partial sums, which the consumer (phase B) adds
// in a fixed order -- deterministic, no atomics.  Every output element sums
// its products in the same order as with 64 x 64 tiles (K step, K slice,
// then hi hi, hi lo, lo hi).
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace pdos_amd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int DT = 64;   // output tile columns
constexpr int DM = 128;  // output tile rows (4 waves of 32)
constexpr int DJ = DT / 16;  // column subtiles a wave covers
constexpr int DK = 64;   // K per step: two MFMA K slices of 32
constexpr int DKP = DK + 8;  // LDS row pitch (bf16) against bank conflicts

// One step's share of a feature-major right-hand side for thread (c, g):
// rows [k + 8 g, k + 8 g + 8) and [k + 32 + 8 g, ...) of column col (rows
// >= rows_valid read as 0), loaded into registers; the loads of the next
// step are issued before the current step is multiplied.
struct ColPart {
  float v[16];
  __device__ __forceinline__ void load(const float* __restrict__ X, int ldb, int rows_valid, int k, int col, int g) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int kk = k + h * 32 + g * 8 + r;
        v[h * 8 + r] = kk < rows_valid ? X[(size_t)kk * ldb + col] : 0.f;
      }
  }
  // split into bf16 hi / lo, one 16-byte LDS store per half and part
  __device__ __forceinline__ void store(__bf16 (*hi)[DKP], __bf16 (*lo)[DKP], int c, int g) const {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      bf16x8 a, b;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const __bf16 hb = (__bf16)v[h * 8 + r];
        a[r] = hb;
        b[r] = (__bf16)(v[h * 8 + r] - (float)hb);
      }
      *reinterpret_cast<bf16x8*>(&hi[c][h * 32 + g * 8]) = a;
      *reinterpret_cast<bf16x8*>(&lo[c][h * 32 + g * 8]) = b;
    }
  }
};

// A operand fragments of one step straight from the split copies (row-major,
// K contiguous, ld elements a row): lane rows fr of the wave's two 16-row
// subtiles, K slices [k + 8 (l / 16), + 8) and the same + 32.
struct AFrag {
  bf16x8 h[2][2], l[2][2];  // [subtile][K slice]
  __device__ __forceinline__ void load(const __bf16* __restrict__ Ah, const __bf16* __restrict__ Al, size_t ld,
                                       int row0, int k, int fr, int fk) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const size_t o = (size_t)(row0 + i * 16 + fr) * ld + k + q * 32 + fk;
        h[i][q] = *reinterpret_cast<const bf16x8*>(Ah + o);
        l[i][q] = *reinterpret_cast<const bf16x8*>(Al + o);
      }
  }
};

// acc += a_hi b_hi + a_hi b_lo + a_lo b_hi over one step (two K slices) for
// the wave's 2 x NJ subtiles of one right-hand side staged in hi / lo.
template <int NJ>
__device__ __forceinline__ void mfma_step(f32x4 (&acc)[2][NJ], const AFrag& A, __bf16 (*hi)[DKP],
                                          __bf16 (*lo)[DKP], int fr, int fk) {
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&hi[j * 16 + fr][q * 32 + fk]);
      const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&lo[j * 16 + fr][q * 32 + fk]);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.h[i][q], bh, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.h[i][q], bl, acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A.l[i][q], bh, acc[i][j], 0, 0, 0);
      }
    }
}

// accumulator lane l holds rows (l / 16) * 4 + r, column l % 16
template <int NJ>
__device__ __forceinline__ void store_tile(const f32x4 (&acc)[2][NJ], float* __restrict__ P, int ldb, int rows,
                                           int i0, int j0, int wm, int lane) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int col = j0 + j * 16 + (lane & 15);
        if (row < rows) P[(size_t)row * ldb + col] = acc[i][j][r];
      }
}

// U (plane z) = A'[i0 : i0 + 128, ks] Y[ks, j0 : j0 + 64], V likewise with W;
// ks = [z kchunk, (z + 1) kchunk) of the rows of A.  ATh / ATl: A' split,
// dense_row_pad(n) x m64 row-major (zeros beyond n, m).  Planes: n x ldb each, plane z at
// z * plane.  Software-pipelined: the next step's A fragments and
// right-hand-side rows are loaded while the current step is multiplied.
__global__ __launch_bounds__(256) void k_dense_at(int m, int n, int ldb, int m64, int kchunk,
                                                  const __bf16* __restrict__ ATh, const __bf16* __restrict__ ATl,
                                                  const float* __restrict__ Y, const float* __restrict__ W,
                                                  float* __restrict__ Up, float* __restrict__ Vp, size_t plane) {
  __shared__ __attribute__((aligned(16))) __bf16 Bs[4][DT][DKP];  // Y hi, Y lo, W hi, W lo
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid * 32;
  const int j0 = blockIdx.x * DT, i0 = blockIdx.y * DM;
  const int k0 = blockIdx.z * kchunk, k1 = min(m64, k0 + kchunk);
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  const int sc = tid & 63, sg = tid >> 6;  // staging: column, group of 8 rows per K slice
  f32x4 acc[2][2][DJ];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < DJ; ++j) acc[h][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  AFrag A, An;
  ColPart y, w;
  if (k0 < k1) {
    A.load(ATh, ATl, m64, i0 + wm, k0, fr, fk);
    y.load(Y, ldb, m, k0, j0 + sc, sg);
    w.load(W, ldb, m, k0, j0 + sc, sg);
  }
  for (int k = k0; k < k1; k += DK) {
    __syncthreads();  // the previous step's B operands are read
    y.store(Bs[0], Bs[1], sc, sg);
    w.store(Bs[2], Bs[3], sc, sg);
    __syncthreads();
    const bool more = k + DK < k1;
    if (more) {
      An.load(ATh, ATl, m64, i0 + wm, k + DK, fr, fk);
      y.load(Y, ldb, m, k + DK, j0 + sc, sg);
      w.load(W, ldb, m, k + DK, j0 + sc, sg);
    }
    mfma_step<DJ>(acc[0], A, Bs[0], Bs[1], fr, fk);
    mfma_step<DJ>(acc[1], A, Bs[2], Bs[3], fr, fk);
    if (more) A = An;
  }
  store_tile<DJ>(acc[0], Up + (size_t)blockIdx.z * plane, ldb, n, i0, j0, wm, lane);
  store_tile<DJ>(acc[1], Vp + (size_t)blockIdx.z * plane, ldb, n, i0, j0, wm, lane);
}

// Z[i0 : i0 + 128, j0 : j0 + 64] = A[i0 : , :] X[:, j0 : ]: ARh / ARl: A
// split, dense_row_pad(m) x n64 row-major (zeros beyond m, n); X: n x ldb.
// Pipelined as k_dense_at.  (128-column tiles, each A fragment used eight
// times, measured slower: 264 registers leave one wave per SIMD.)
__global__ __launch_bounds__(256) void k_dense_a(int m, int n, int ldb, int n64, const __bf16* __restrict__ ARh,
                                                 const __bf16* __restrict__ ARl, const float* __restrict__ X,
                                                 float* __restrict__ Z) {
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][DT][DKP];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid * 32;
  const int j0 = blockIdx.x * DT, i0 = blockIdx.y * DM;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  const int sc = tid & 63, sg = tid >> 6;
  f32x4 acc[2][DJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < DJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  AFrag A, An;
  ColPart x;
  A.load(ARh, ARl, n64, i0 + wm, 0, fr, fk);
  x.load(X, ldb, n, 0, j0 + sc, sg);
  for (int k = 0; k < n64; k += DK) {
    __syncthreads();
    x.store(Bs[0], Bs[1], sc, sg);
    __syncthreads();
    const bool more = k + DK < n64;
    if (more) {
      An.load(ARh, ARl, n64, i0 + wm, k + DK, fr, fk);
      x.load(X, ldb, n, k + DK, j0 + sc, sg);
    }
    mfma_step<DJ>(acc, A, Bs[0], Bs[1], fr, fk);
    if (more) A = An;
  }
  store_tile<DJ>(acc, Z, ldb, m, i0, j0, wm, lane);
}

}  // namespace

int dense_row_pad(int rows) { return (rows + DM - 1) / DM * DM; }

int dense_at_splits(int n, int m64, int ldb) {
  const int tiles = ((n + DM - 1) / DM) * (ldb / DT);
  int s = 1;
  while (s < 16 && tiles * s * 2 <= 512 && m64 / (s * 2) >= 4 * DK) s *= 2;
  return s;
}

hipError_t dense_at_pair(hipStream_t st, int m, int n, int ldb, int m64, const __bf16* ATh, const __bf16* ATl,
                         const float* Y, const float* W, float* Up, float* Vp, size_t plane, int splits) {
  if (ldb % DT || m64 % DK || splits < 1) return hipErrorInvalidValue;
  const int kchunk = ((m64 / splits + DK - 1) / DK) * DK;
  hipLaunchKernelGGL(k_dense_at, dim3(ldb / DT, (n + DM - 1) / DM, splits), dim3(256), 0, st, m, n, ldb, m64,
                     kchunk, ATh, ATl, Y, W, Up, Vp, plane);
  return hipGetLastError();
}

hipError_t dense_a(hipStream_t st, int m, int n, int ldb, int n64, const __bf16* ARh, const __bf16* ARl,
                   const float* X, float* Z) {
  if (ldb % DT || n64 % DK) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dense_a, dim3(ldb / DT, (m + DM - 1) / DM), dim3(256), 0, st, m, n, ldb, n64, ARh, ARl, X, Z);
  return hipGetLastError();
}

}  // namespace pdos_amd
