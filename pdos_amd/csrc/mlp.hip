// mlp.hip -- the learned warm start: a 3-layer MLP (in -> H -> H -> out, GELU)
// trained and evaluated with hand-written CDNA4 kernels.
//
// Weights are kept in binary32 (master copy, Adam state) and mirrored in
// bfloat16 for the products; every product is one BF16 MFMA GEMM
// (v_mfma_f32_16x16x32_bf16, binary32 accumulation) with a fused epilogue:
//
//   forward   Z1 = X W1 + b1, H1 = gelu(Z1)      [bias+GELU epilogue]
//             Z2 = H1 W2 + b2, H2 = gelu(Z2)
//             Y  = H2 W3 + b3                    [bias epilogue]
//   backward  dY = 2 (Y - T) / count (masked to valid rows / columns)
//             dW3 = H2' dY, db3 = colsum dY      [A-transposed GEMM]
//             dZ2 = (dY W3') .* gelu'(Z2)         [B-transposed GEMM, GELU' epilogue]
//             dW2 = H1' dZ2, dZ1 = (dZ2 W2') .* gelu'(Z1), dW1 = X' dZ1
//   update    Adam on the binary32 master, then refresh the bf16 mirror
//
// All dimensions are padded to multiples of 64 (zero rows/columns, masked in
// the loss), so the GEMM tiles never need bounds checks.  Inputs and outputs
// pass through affine maps (standardisation) folded into the first layer's
// input kernel and the last layer's epilogue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/pdos_ext.h"

namespace pdos_amd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 64;  // K step of the MLP products (the family's x-update uses 128)
constexpr int GT = 256;

enum Epi { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_AFFINE = 2, EPI_GELU_GRAD = 3, EPI_ADD = 4 };

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C[M x N] = op(A) op(B) with bf16 operands and fp32 accumulation.
//   TA = false: A is M x K row-major (lda);  TA = true: A is stored K x M (lda)
//   TB = false: B is K x N row-major (ldb);  TB = true: B is stored N x K (ldb)
// M % 64 == 0, N % 64 == 0, K % KB == 0 (KB = 64 or 128).  Block: 64 x 64
// tile, 4 waves in a 2 x 2 layout, each wave 32 x 32 = 2 x 2 MFMA tiles of
// 16 x 16.  The K loop is software-pipelined two steps deep: two register
// sets alternate, and the global loads of steps k + 1 and k + 2 are in
// flight while step k is multiplied out of LDS.
template <bool TA, bool TB, int EPI, int KB = BK>
__global__ __launch_bounds__(GT) void k_gemm(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                             const __bf16* __restrict__ B, int ldb, float* __restrict__ Cf,
                                             __bf16* __restrict__ Cb, int ldc, const float* __restrict__ bias,
                                             const float* __restrict__ aux, const float* __restrict__ omul,
                                             const float* __restrict__ oadd) {
  constexpr int KP = KB + 8;  // LDS rows padded against bank conflicts
  constexpr int H = KB / 32;   // 16-byte pieces per thread, tile and K step
  __shared__ __attribute__((aligned(16))) __bf16 As[BM][KP];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[BN][KP];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Each thread moves H 16-byte pieces of each tile per K step.
  //   row-major source (A: !TA, B: TB):  row t / 4, k (t % 4) * 8 + 32 h
  //   transposed source (A: TA, B: !TB): k t / 8 + 32 h, row (t % 8) * 8 .. + 7
  bf16x8 ra0[H], rb0[H], ra1[H], rb1[H];
  auto gload = [&](int k0, bf16x8 (&ra)[H], bf16x8 (&rb)[H]) {
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (!TA) ra[h] = *reinterpret_cast<const bf16x8*>(&A[(size_t)(m0 + (tid >> 2)) * lda + k0 + (tid & 3) * 8 + h * 32]);
      else ra[h] = *reinterpret_cast<const bf16x8*>(&A[(size_t)(k0 + (tid >> 3) + h * 32) * lda + m0 + (tid & 7) * 8]);
      if (TB) rb[h] = *reinterpret_cast<const bf16x8*>(&B[(size_t)(n0 + (tid >> 2)) * ldb + k0 + (tid & 3) * 8 + h * 32]);
      else rb[h] = *reinterpret_cast<const bf16x8*>(&B[(size_t)(k0 + (tid >> 3) + h * 32) * ldb + n0 + (tid & 7) * 8]);
    }
  };
  auto sstore = [&](const bf16x8 (&ra)[H], const bf16x8 (&rb)[H]) {
#pragma unroll
    for (int h = 0; h < H; ++h) {
      if (!TA) {
        *reinterpret_cast<bf16x8*>(&As[tid >> 2][(tid & 3) * 8 + h * 32]) = ra[h];
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) As[(tid & 7) * 8 + i][(tid >> 3) + h * 32] = ra[h][i];
      }
      if (TB) {
        *reinterpret_cast<bf16x8*>(&Bs[tid >> 2][(tid & 3) * 8 + h * 32]) = rb[h];
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) Bs[(tid & 7) * 8 + i][(tid >> 3) + h * 32] = rb[h][i];
      }
    }
  };
  gload(0, ra0, rb0);
  if (KB < K) gload(KB, ra1, rb1);
  // EPI_ADD: the addend tile, loaded now so it arrives during the K loop
  float addv[2][2][4];
  if (EPI == EPI_ADD) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
          const int col = n0 + wn + j * 16 + (lane & 15);
          addv[i][j][r] = aux[(size_t)row * ldc + col];
        }
  }
  // lane l: A rows (l % 16), k = (l / 16) * 8 .. +7; B cols (l % 16), same k
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  auto multiply = [&]() {
#pragma unroll
    for (int ks = 0; ks < KB; ks += 32) {
      bf16x8 af[2], bfv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8*>(&As[wm + i * 16 + fr][ks + fk]);
#pragma unroll
      for (int j = 0; j < 2; ++j) bfv[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + j * 16 + fr][ks + fk]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfv[j], acc[i][j], 0, 0, 0);
    }
  };
  // steps k0 (set 0) and k0 + KB (set 1); a set is refilled with the step
  // two ahead as soon as it is in LDS
  for (int k0 = 0; k0 < K; k0 += 2 * KB) {
    sstore(ra0, rb0);
    __syncthreads();
    if (k0 + 2 * KB < K) gload(k0 + 2 * KB, ra0, rb0);
    multiply();
    __syncthreads();
    if (k0 + KB >= K) break;
    sstore(ra1, rb1);
    __syncthreads();
    if (k0 + 3 * KB < K) gload(k0 + 3 * KB, ra1, rb1);
    multiply();
    __syncthreads();
  }
  // accumulator lane l holds rows (l / 16) * 4 + r, column l % 16
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int col = n0 + wn + j * 16 + (lane & 15);
        const size_t o = (size_t)row * ldc + col;
        float v = acc[i][j][r];
        if (EPI == EPI_F32) {
          Cf[o] = v;
        } else if (EPI == EPI_ADD) {  // aux is an addend with the layout of C
          Cf[o] = addv[i][j][r] + v;
        } else if (EPI == EPI_BIAS_GELU) {
          v += bias[col];
          Cf[o] = v;                 // pre-activation, kept for the backward pass
          Cb[o] = (__bf16)gelu(v);   // activation, the next layer's operand
        } else if (EPI == EPI_BIAS_AFFINE) {
          v += bias[col];
          Cf[o] = omul ? v * omul[col] + oadd[col] : v;
        } else {  // EPI_GELU_GRAD: v is dH; aux holds Z
          Cb[o] = (__bf16)(v * gelu_grad(aux[o]));
        }
      }
}

// Standardised, zero-padded bf16 input rows: X'[r][c] = (X[r][c] - mu[c]) / sd[c].
__global__ void k_prep_input(int rows, int cols, int ld_in, const float* __restrict__ X,
                             const float* __restrict__ mu, const float* __restrict__ sd, int ld_out,
                             int rows_pad, __bf16* __restrict__ Xb) {
  const long total = (long)rows_pad * ld_out;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e / ld_out), c = (int)(e % ld_out);
    float v = 0.f;
    if (r < rows && c < cols) v = (X[(size_t)r * ld_in + c] - mu[c]) / sd[c];
    Xb[e] = (__bf16)v;
  }
}

// dY = 2 (Y - T) / (rows * cols) on valid entries, 0 on padding; per-block
// partial sums of (Y - T)^2.  T is standardised (cols valid of ld).
__global__ void k_loss_grad(int rows, int cols, int ld, const float* __restrict__ Y,
                            const float* __restrict__ T, int ldt, int rows_pad, float* __restrict__ dYf,
                            __bf16* __restrict__ dYb, double* __restrict__ partial) {
  __shared__ double red[GT / 64];
  const long total = (long)rows_pad * ld;
  const float scale = 2.0f / (float)((double)rows * cols);
  double acc = 0.0;
  // (r, c) of e, advanced by the stride's row and column steps (no 64-bit
  // division per element)
  const long e0 = blockIdx.x * (long)blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
  int r = (int)(e0 / ld), c = (int)(e0 % ld);
  const int sr = (int)(stride / ld), sc = (int)(stride % ld);
  for (long e = e0; e < total; e += stride) {
    float g = 0.f;
    if (r < rows && c < cols) {
      const float d = Y[e] - T[(size_t)r * ldt + c];
      acc += (double)d * d;
      g = scale * d;
    }
    dYf[e] = g;
    dYb[e] = (__bf16)g;
    r += sr;
    c += sc;
    if (c >= ld) {
      c -= ld;
      ++r;
    }
  }
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < GT / 64; ++w) s += red[w];
    partial[blockIdx.x] = s;
  }
}

// db[c] = sum_r G[r][c] over rows (fp32 or bf16 source).  A block of 512
// threads takes 64 columns (lane = column, coalesced rows); wave w sums rows
// w, w + 8, ... and the 8 wave sums are added in wave order (deterministic).
template <typename S>
__device__ __forceinline__ void colsum_block(int blk, int rows, int cols, int ld, const S* __restrict__ G,
                                             float* __restrict__ db) {
  __shared__ float part[8][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = blk * 64 + lane;
  float s = 0.f;
  if (c < cols) {
#pragma unroll 8
    for (int r = w; r < rows; r += 8) s += (float)G[(size_t)r * ld + c];
  }
  part[w][lane] = s;
  __syncthreads();
  if (w == 0 && c < cols) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) t += part[k][lane];
    db[c] = t;
  }
}

// The three bias gradients of the backward pass in one launch: blocks
// [0, n3) colsum dY (fp32) into db3, then [n3, n3 + nh) dZ2 (bf16) into db2,
// then [n3 + nh, ..) dZ1 into db1 -- each column block as colsum_block.
__global__ __launch_bounds__(512) void k_colsum3(int rows, int pout, int ph, const float* __restrict__ dY,
                                                 const __bf16* __restrict__ dZ2, const __bf16* __restrict__ dZ1,
                                                 float* __restrict__ db3, float* __restrict__ db2,
                                                 float* __restrict__ db1) {
  const int n3 = (pout + 63) / 64, nh = (ph + 63) / 64;
  const int b = blockIdx.x;
  if (b < n3) colsum_block<float>(b, rows, pout, pout, dY, db3);
  else if (b < n3 + nh) colsum_block<__bf16>(b - n3, rows, ph, ph, dZ2, db2);
  else colsum_block<__bf16>(b - n3 - nh, rows, ph, ph, dZ1, db1);
}

// The bf16 mirrors of the weight matrices: master entries [o[k], o[k] + len[k])
// are mirrored at p[k].
struct BfMirror {
  long o[3], len[3];
  __bf16* p[3];
};

// Adam (binary32 master, bias-corrected), with the updated weights written to
// their bf16 mirrors.
__global__ void k_adam(long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, BfMirror mir, float lr, float b1, float b2, float c1, float c2,
                       float eps) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if ((unsigned long)(i - mir.o[k]) < (unsigned long)mir.len[k]) mir.p[k][i - mir.o[k]] = (__bf16)wi;
  }
}

// A training step's minibatch: Xb[r][c] = bf16((X[idx[r]][c] - mu[c]) / sd[c])
// over the padded rows_pad x ld_out (zero outside rows x nin; the arithmetic
// of k_prep_input on the gathered rows), and Td[r][:] = Ts[idx[r]][:].
__global__ void k_gather_prep(int rows, int nin, int ld_out, int rows_pad, int nout, const int* __restrict__ idx,
                              const float* __restrict__ X, const float* __restrict__ mu,
                              const float* __restrict__ sd, __bf16* __restrict__ Xb, const float* __restrict__ Ts,
                              float* __restrict__ Td) {
  const long stride = (long)gridDim.x * blockDim.x, e0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
  {
    const long total = (long)rows_pad * ld_out;
    int r = (int)(e0 / ld_out), c = (int)(e0 % ld_out);
    const int sr = (int)(stride / ld_out), sc = (int)(stride % ld_out);
    for (long e = e0; e < total; e += stride) {
      float v = 0.f;
      if (r < rows && c < nin) v = (X[(size_t)idx[r] * nin + c] - mu[c]) / sd[c];
      Xb[e] = (__bf16)v;
      r += sr;
      c += sc;
      if (c >= ld_out) {
        c -= ld_out;
        ++r;
      }
    }
  }
  const long total = (long)rows * nout;
  int r = (int)(e0 / nout), c = (int)(e0 % nout);
  const int sr = (int)(stride / nout), sc = (int)(stride % nout);
  for (long e = e0; e < total; e += stride) {
    Td[e] = Ts[(size_t)idx[r] * nout + c];
    r += sr;
    c += sc;
    if (c >= nout) {
      c -= nout;
      ++r;
    }
  }
}

__global__ void k_to_bf16(long n, const float* __restrict__ a, __bf16* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (__bf16)a[i];
}

int pad64(int v) { return (v + 63) & ~63; }

// Launches k_gemm; shapes off the 64-multiple grid are refused (the kernel
// has no bounds checks).
template <bool TA, bool TB, int EPI, int KB = BK>
hipError_t gemm(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
                float* Cf, __bf16* Cb, int ldc, const float* bias = nullptr, const float* aux = nullptr,
                const float* omul = nullptr, const float* oadd = nullptr) {
  if (M <= 0 || N <= 0 || K <= 0 || M % BM || N % BN || K % KB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI, KB>), dim3(N / BN, M / BM), dim3(GT), 0, st, M, N, K, A, lda, B, ldb,
                     Cf, Cb, ldc, bias, aux, omul, oadd);
  return hipGetLastError();
}

// gemm with 128-wide K steps where K allows (half the dependent load round
// trips of the 64-wide steps; the MLP products are latency-bound tiles).
template <bool TA, bool TB, int EPI>
hipError_t gemm_k(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
                  float* Cf, __bf16* Cb, int ldc, const float* bias = nullptr, const float* aux = nullptr,
                  const float* omul = nullptr, const float* oadd = nullptr) {
  if (K % 128 == 0) return gemm<TA, TB, EPI, 128>(st, M, N, K, A, lda, B, ldb, Cf, Cb, ldc, bias, aux, omul, oadd);
  return gemm<TA, TB, EPI>(st, M, N, K, A, lda, B, ldb, Cf, Cb, ldc, bias, aux, omul, oadd);
}

}  // namespace

// C (fp32, M x N) = A (bf16, M x K) B (bf16, K x N), row-major; M, N and K
// multiples of 64; with `addend` (M x N, ldc), C = addend + A B.  Used by
// the family engine's x-update.
hipError_t bf16_gemm_f32(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* B,
                         int ldb, float* C, int ldc, const float* addend) {
  // 128-wide K steps when K allows: half the dependent load round trips
  if (K % 128 == 0) {
    if (addend)
      return gemm<false, false, EPI_ADD, 128>(st, M, N, K, A, lda, B, ldb, C, nullptr, ldc, nullptr, addend);
    return gemm<false, false, EPI_F32, 128>(st, M, N, K, A, lda, B, ldb, C, nullptr, ldc);
  }
  if (addend) return gemm<false, false, EPI_ADD>(st, M, N, K, A, lda, B, ldb, C, nullptr, ldc, nullptr, addend);
  return gemm<false, false, EPI_F32>(st, M, N, K, A, lda, B, ldb, C, nullptr, ldc);
}

// ---------------------------------------------------------------------------

struct Mlp {
  int device = 0;
  int nin = 0, nh = 0, nout = 0;        // logical sizes
  int pin = 0, ph = 0, pout = 0;        // padded sizes
  int cap = 0;                          // padded rows the work buffers hold
  hipStream_t st = nullptr;
  // parameters: fp32 master, Adam state, bf16 mirror (weights only)
  float *W1 = nullptr, *b1 = nullptr, *W2 = nullptr, *b2 = nullptr, *W3 = nullptr, *b3 = nullptr;
  float *g = nullptr, *mom = nullptr, *vel = nullptr;  // one block each, same layout as the master
  float* master = nullptr;
  __bf16 *W1b = nullptr, *W2b = nullptr, *W3b = nullptr;
  float *in_mu = nullptr, *in_sd = nullptr, *out_mul = nullptr, *out_add = nullptr;
  // activations
  __bf16 *Xb = nullptr, *H1b = nullptr, *H2b = nullptr, *dYb = nullptr, *dZ2b = nullptr, *dZ1b = nullptr;
  float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *dYf = nullptr, *T = nullptr, *Xf = nullptr;
  double* part = nullptr;
  long step = 0;
  long nparam = 0;
  size_t off_W1 = 0, off_b1 = 0, off_W2 = 0, off_b2 = 0, off_W3 = 0, off_b3 = 0;
  std::vector<void*> allocs;

  template <typename P>
  bool alloc(P*& p, size_t count, std::string& err) {
    void* q = nullptr;
    if (hipMalloc(&q, count * sizeof(P) + 256) != hipSuccess) { err = "hipMalloc failed"; return false; }
    (void)hipMemset(q, 0, count * sizeof(P) + 256);
    allocs.push_back(q);
    p = static_cast<P*>(q);
    return true;
  }
  ~Mlp() {
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  bool ensure_rows(int rows, std::string& err) {
    const int need = pad64(rows);
    if (need <= cap) return true;
    cap = need;
    return alloc(Xb, (size_t)cap * pin, err) && alloc(H1b, (size_t)cap * ph, err) &&
           alloc(H2b, (size_t)cap * ph, err) && alloc(dYb, (size_t)cap * pout, err) &&
           alloc(dZ2b, (size_t)cap * ph, err) && alloc(dZ1b, (size_t)cap * ph, err) &&
           alloc(Z1, (size_t)cap * ph, err) && alloc(Z2, (size_t)cap * ph, err) &&
           alloc(Y, (size_t)cap * pout, err) && alloc(dYf, (size_t)cap * pout, err) &&
           alloc(T, (size_t)cap * nout, err) && alloc(Xf, (size_t)cap * nin, err) &&
           alloc(part, 1024, err);
  }
  // first refused GEMM launch since the last API call (none expected: every
  // dimension is padded to a multiple of 64)
  hipError_t gerr = hipSuccess;
  void note(hipError_t e) {
    if (gerr == hipSuccess && e != hipSuccess) gerr = e;
  }
  void refresh_bf16() {
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, st, (long)pin * ph, W1, W1b);
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, st, (long)ph * ph, W2, W2b);
    hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, st, (long)ph * pout, W3, W3b);
  }
  // forward for `rows` rows whose raw inputs are in Xf (device, rows x nin)
  void forward(int rows, bool affine_out) { forward_into(rows, affine_out, Xf, Xb, Z1, H1b, Z2, H2b, Y); }
  // the same with explicit activation buffers (pad64(rows) rows each): X raw
  // inputs (rows x nin), xb standardised input, z1 / z2 pre-activations (may
  // alias each other: the forward pass never reads them), h1 / h2 hidden
  // layers, y the padded output
  void forward_into(int rows, bool affine_out, const float* X, __bf16* xb, float* z1, __bf16* h1, float* z2,
                    __bf16* h2, float* y) {
    const int rp = pad64(rows);
    hipLaunchKernelGGL(k_prep_input, dim3(512), dim3(256), 0, st, rows, nin, nin, X, in_mu, in_sd, pin, rp, xb);
    forward_layers(rows, affine_out, xb, z1, h1, z2, h2, y);
  }
  // the three products from the standardized bf16 input xb
  void forward_layers(int rows, bool affine_out, const __bf16* xb, float* z1, __bf16* h1, float* z2, __bf16* h2,
                      float* y) {
    const int rp = pad64(rows);
    note(gemm_k<false, false, EPI_BIAS_GELU>(st, rp, ph, pin, xb, pin, W1b, ph, z1, h1, ph, b1));
    note(gemm_k<false, false, EPI_BIAS_GELU>(st, rp, ph, ph, h1, ph, W2b, ph, z2, h2, ph, b2));
    note(gemm_k<false, false, EPI_BIAS_AFFINE>(st, rp, pout, ph, h2, ph, W3b, pout, y, nullptr, pout, b3, nullptr,
                                        affine_out ? out_mul : nullptr, affine_out ? out_add : nullptr));
  }
  // loss and parameter gradients into g for the rows in Xf / T; the loss
  // (mean square error) is returned after one synchronization
  static constexpr int kLossBlocks = 256;
  double backward(int rows) {
    backward_async(rows, part);
    std::vector<double> hp(kLossBlocks);
    (void)hipMemcpyAsync(hp.data(), part, kLossBlocks * sizeof(double), hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    return loss_of(hp.data(), rows);
  }
  // the loss from k_loss_grad's per-block partials, summed in block order
  double loss_of(const double* hp, int rows) const {
    double s = 0.0;
    for (int v = 0; v < kLossBlocks; ++v) s += hp[v];
    return s / ((double)rows * nout);
  }
  // the same without waiting: the loss partials go to `partial`
  // (kLossBlocks doubles), the gradients into g
  void backward_async(int rows, double* partial) {
    const int rp = pad64(rows);
    const int nb = kLossBlocks;
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(GT), 0, st, rows, nout, pout, Y, T, nout, rp, dYf, dYb, partial);
    float* gW1 = g + off_W1; float* gb1 = g + off_b1;
    float* gW2 = g + off_W2; float* gb2 = g + off_b2;
    float* gW3 = g + off_W3; float* gb3 = g + off_b3;
    note(gemm_k<true, false, EPI_F32>(st, ph, pout, rp, H2b, ph, dYb, pout, gW3, nullptr, pout));
    note(gemm_k<false, true, EPI_GELU_GRAD>(st, rp, ph, pout, dYb, pout, W3b, pout, nullptr, dZ2b, ph, nullptr, Z2));
    note(gemm_k<true, false, EPI_F32>(st, ph, ph, rp, H1b, ph, dZ2b, ph, gW2, nullptr, ph));
    note(gemm_k<false, true, EPI_GELU_GRAD>(st, rp, ph, ph, dZ2b, ph, W2b, ph, nullptr, dZ1b, ph, nullptr, Z1));
    note(gemm_k<true, false, EPI_F32>(st, pin, ph, rp, Xb, pin, dZ1b, ph, gW1, nullptr, ph));
    hipLaunchKernelGGL(k_colsum3, dim3((pout + 63) / 64 + 2 * ((ph + 63) / 64)), dim3(512), 0, st, rp, pout, ph, dYf,
                       dZ2b, dZ1b, gb3, gb2, gb1);
  }
  void adam(float lr) {
    ++step;
    const float b1c = 0.9f, b2c = 0.999f;
    const float c1 = 1.f - std::pow(b1c, (float)step), c2 = 1.f - std::pow(b2c, (float)step);
    BfMirror mir;
    mir.o[0] = (long)off_W1, mir.len[0] = (long)pin * ph, mir.p[0] = W1b;
    mir.o[1] = (long)off_W2, mir.len[1] = (long)ph * ph, mir.p[1] = W2b;
    mir.o[2] = (long)off_W3, mir.len[2] = (long)ph * pout, mir.p[2] = W3b;
    hipLaunchKernelGGL(k_adam, dim3(1024), dim3(256), 0, st, nparam, master, g, mom, vel, mir, lr, b1c, b2c, c1,
                       c2, 1e-8f);
  }
};

}  // namespace pdos_amd

using pdos_amd::Mlp;

// Scratch of pdos_mlp_predict_dev_noalloc: the forward pass's activations for
// one chunk of rows (standardised bf16 input, one pre-activation buffer
// shared by both hidden layers, the two bf16 hidden layers, the padded
// binary32 output).  One copy per GPU the library runs on, held until the
// process ends; every no-alloc call on a GPU shares it.
constexpr size_t kMlpScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kMlpScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 256;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};        // device address of `scratch` per GPU
hipError_t g_scratch_err[kMaxScratchDevices] = {};

// Once per device, with that device current: the device address of this
// GPU's copy of `scratch` (the host-side name is not a device address).
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_dev[device] = p;
}

unsigned char* scratch_of(int device, std::string& err) {
  if (device < 0 || device >= kMaxScratchDevices) { err = "device ordinal out of range"; return nullptr; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    err = std::string("scratch not set up on this device: hipGetSymbolAddress: ") +
          hipGetErrorString(g_scratch_err[device]);
    return nullptr;
  }
  return static_cast<unsigned char*>(g_scratch_dev[device]);
}
}  // namespace

struct PdosMlp {
  Mlp net;
};

namespace {
thread_local std::string g_mlp_err;
}

extern "C" {

const char* pdos_mlp_last_error(void) { return g_mlp_err.c_str(); }

PdosMlp* pdos_mlp_create(int n_in, int n_hidden, int n_out, unsigned long seed, int device) {
  using namespace pdos_amd;
  if (n_in < 1 || n_hidden < 1 || n_out < 1) { g_mlp_err = "bad sizes"; return nullptr; }
  if (hipSetDevice(device) != hipSuccess) { g_mlp_err = "hipSetDevice failed"; return nullptr; }
  scratch_setup(device);
  PdosMlp* P = new PdosMlp;
  Mlp& M = P->net;
  std::string err;
  M.device = device;
  M.nin = n_in; M.nh = n_hidden; M.nout = n_out;
  M.pin = pad64(n_in); M.ph = pad64(n_hidden); M.pout = pad64(n_out);
  M.off_W1 = 0;
  M.off_b1 = M.off_W1 + (size_t)M.pin * M.ph;
  M.off_W2 = M.off_b1 + M.ph;
  M.off_b2 = M.off_W2 + (size_t)M.ph * M.ph;
  M.off_W3 = M.off_b2 + M.ph;
  M.off_b3 = M.off_W3 + (size_t)M.ph * M.pout;
  M.nparam = (long)(M.off_b3 + M.pout);
  if (hipStreamCreateWithFlags(&M.st, hipStreamNonBlocking) != hipSuccess ||
      !M.alloc(M.master, M.nparam, err) || !M.alloc(M.g, M.nparam, err) || !M.alloc(M.mom, M.nparam, err) ||
      !M.alloc(M.vel, M.nparam, err) || !M.alloc(M.W1b, (size_t)M.pin * M.ph, err) ||
      !M.alloc(M.W2b, (size_t)M.ph * M.ph, err) || !M.alloc(M.W3b, (size_t)M.ph * M.pout, err) ||
      !M.alloc(M.in_mu, M.nin, err) || !M.alloc(M.in_sd, M.nin, err) || !M.alloc(M.out_mul, M.pout, err) ||
      !M.alloc(M.out_add, M.pout, err)) {
    g_mlp_err = err.empty() ? "stream create failed" : err;
    delete P;
    return nullptr;
  }
  M.W1 = M.master + M.off_W1; M.b1 = M.master + M.off_b1;
  M.W2 = M.master + M.off_W2; M.b2 = M.master + M.off_b2;
  M.W3 = M.master + M.off_W3; M.b3 = M.master + M.off_b3;
  // He-style init on the logical block, zeros on the padding
  std::vector<float> h(M.nparam, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](size_t off, int rows, int cols, int ld, float scale) {
    std::normal_distribution<float> nd(0.f, scale);
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c) h[off + (size_t)r * ld + c] = nd(rng);
  };
  fill(M.off_W1, M.nin, M.nh, M.ph, std::sqrt(2.f / M.nin));
  fill(M.off_W2, M.nh, M.nh, M.ph, std::sqrt(2.f / M.nh));
  fill(M.off_W3, M.nh, M.nout, M.pout, std::sqrt(1.f / M.nh));
  std::vector<float> ones(M.pout, 1.f), zeros(M.pout, 0.f);
  (void)hipMemcpy(M.master, h.data(), h.size() * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(M.in_mu, zeros.data(), M.nin * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(M.in_sd, ones.data(), M.nin * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(M.out_mul, ones.data(), M.pout * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(M.out_add, zeros.data(), M.pout * 4, hipMemcpyHostToDevice);
  M.refresh_bf16();
  (void)hipStreamSynchronize(M.st);
  return P;
}

void pdos_mlp_free(PdosMlp* P) { delete P; }

int pdos_mlp_num_params(const PdosMlp* P) {
  const Mlp& M = P->net;
  return M.nin * M.nh + M.nh + M.nh * M.nh + M.nh + M.nh * M.nout + M.nout;
}

// Parameters in logical layout W1 (in x h), b1, W2 (h x h), b2, W3 (h x out), b3.
static void copy_params(Mlp& M, float* host, bool to_device) {
  size_t o = 0;
  auto blk = [&](size_t off, int rows, int cols, int ld) {
    for (int r = 0; r < rows; ++r) {
      float* dev = M.master + off + (size_t)r * ld;
      if (to_device) (void)hipMemcpy(dev, host + o, cols * 4, hipMemcpyHostToDevice);
      else (void)hipMemcpy(host + o, dev, cols * 4, hipMemcpyDeviceToHost);
      o += cols;
    }
  };
  blk(M.off_W1, M.nin, M.nh, M.ph);
  blk(M.off_b1, 1, M.nh, M.ph);
  blk(M.off_W2, M.nh, M.nh, M.ph);
  blk(M.off_b2, 1, M.nh, M.ph);
  blk(M.off_W3, M.nh, M.nout, M.pout);
  blk(M.off_b3, 1, M.nout, M.pout);
}

int pdos_mlp_get_params(PdosMlp* P, float* out) {
  copy_params(P->net, out, false);
  return 0;
}

int pdos_mlp_set_params(PdosMlp* P, const float* in) {
  copy_params(P->net, const_cast<float*>(in), true);
  P->net.refresh_bf16();
  return hipStreamSynchronize(P->net.st) == hipSuccess ? 0 : -1;
}

int pdos_mlp_set_affine(PdosMlp* P, const float* in_mean, const float* in_std, const float* out_std,
                        const float* out_mean) {
  Mlp& M = P->net;
  if (in_mean) (void)hipMemcpy(M.in_mu, in_mean, M.nin * 4, hipMemcpyHostToDevice);
  if (in_std) (void)hipMemcpy(M.in_sd, in_std, M.nin * 4, hipMemcpyHostToDevice);
  if (out_std) (void)hipMemcpy(M.out_mul, out_std, M.nout * 4, hipMemcpyHostToDevice);
  if (out_mean) (void)hipMemcpy(M.out_add, out_mean, M.nout * 4, hipMemcpyHostToDevice);
  return 0;
}

// 0 when the stream is idle without error and no GEMM launch was refused.
static int mlp_status(Mlp& M) {
  const bool ok = hipStreamSynchronize(M.st) == hipSuccess && M.gerr == hipSuccess;
  if (M.gerr != hipSuccess) g_mlp_err = "GEMM shape off the 64-multiple grid";
  M.gerr = hipSuccess;
  return ok ? 0 : -1;
}

// Loss and gradients (logical layout, same as get_params) of one batch;
// targets T are in the standardised output space.
int pdos_mlp_loss_grad(PdosMlp* P, const float* X, const float* T, int rows, double* loss, float* grads) {
  Mlp& M = P->net;
  std::string err;
  if (!M.ensure_rows(rows, err)) { g_mlp_err = err; return -1; }
  (void)hipMemcpy(M.Xf, X, (size_t)rows * M.nin * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(M.T, T, (size_t)rows * M.nout * 4, hipMemcpyHostToDevice);
  M.forward(rows, false);
  const double l = M.backward(rows);
  if (loss) *loss = l;
  if (grads) {
    float* save = M.master;
    M.master = M.g;  // reuse the layout walker on the gradient block
    copy_params(M, grads, false);
    M.master = save;
  }
  return mlp_status(M);
}

// Trains `epochs` passes over (X, T) (host arrays, rows x nin / rows x nout,
// T standardised) in shuffled minibatches; returns the last epoch's mean loss.
int pdos_mlp_train(PdosMlp* P, const float* X, const float* T, int rows, int epochs, int batch, float lr,
                   unsigned long seed, double* last_loss) {
  Mlp& M = P->net;
  std::string err;
  if (batch < 1 || rows < 1) { g_mlp_err = "bad batch"; return -1; }
  if (!M.ensure_rows(batch, err)) { g_mlp_err = err; return -1; }
  // the whole training set lives on the device; minibatches are gathered by index
  float *dX = nullptr, *dT = nullptr;
  if (hipMalloc(&dX, (size_t)rows * M.nin * 4) != hipSuccess ||
      hipMalloc(&dT, (size_t)rows * M.nout * 4) != hipSuccess) {
    g_mlp_err = "hipMalloc failed";
    return -1;
  }
  (void)hipMemcpy(dX, X, (size_t)rows * M.nin * 4, hipMemcpyHostToDevice);
  (void)hipMemcpy(dT, T, (size_t)rows * M.nout * 4, hipMemcpyHostToDevice);
  std::vector<int> order(rows);
  for (int i = 0; i < rows; ++i) order[i] = i;
  int* dIdx = nullptr;
  if (hipMalloc(&dIdx, (size_t)rows * 4) != hipSuccess) { g_mlp_err = "hipMalloc failed"; return -1; }
  // An epoch's steps are enqueued without waiting: each step's loss
  // partials go to its own slot, and the host waits once per epoch (before
  // the next epoch's indices are copied from the buffer it shuffles into
  // while the GPU works); the last epoch's partials give the loss returned,
  // summed as a per-step wait would have.
  const int steps = rows / batch;
  double* dPart = nullptr;
  if (hipMalloc(&dPart, (size_t)std::max(steps, 1) * Mlp::kLossBlocks * sizeof(double)) != hipSuccess) {
    g_mlp_err = "hipMalloc failed";
    return -1;
  }
  std::vector<int> order2(order);
  std::vector<int>* ord[2] = {&order, &order2};
  std::mt19937_64 rng(seed);
  std::shuffle(order.begin(), order.end(), rng);
  double epoch_loss = 0.0;
  for (int e = 0; e < epochs; ++e) {
    std::vector<int>& cur = *ord[e & 1];
    (void)hipMemcpyAsync(dIdx, cur.data(), (size_t)rows * 4, hipMemcpyHostToDevice, M.st);
    for (int q = 0; q < steps; ++q) {
      const int s = q * batch;
      hipLaunchKernelGGL(pdos_amd::k_gather_prep, dim3(1024), dim3(256), 0, M.st, batch, M.nin, M.pin,
                         pdos_amd::pad64(batch), M.nout, dIdx + s, dX, M.in_mu, M.in_sd, M.Xb, dT, M.T);
      M.forward_layers(batch, false, M.Xb, M.Z1, M.H1b, M.Z2, M.H2b, M.Y);
      M.backward_async(batch, dPart + (size_t)q * Mlp::kLossBlocks);
      M.adam(lr);
    }
    if (e + 1 < epochs) {  // the next epoch's order, shuffled from this one's as in place
      std::vector<int>& nxt = *ord[(e + 1) & 1];
      nxt = cur;
      std::shuffle(nxt.begin(), nxt.end(), rng);
    }
    (void)hipStreamSynchronize(M.st);
  }
  if (epochs > 0 && steps > 0) {
    std::vector<double> hp((size_t)steps * Mlp::kLossBlocks);
    (void)hipMemcpy(hp.data(), dPart, hp.size() * sizeof(double), hipMemcpyDeviceToHost);
    double acc = 0.0;
    for (int q = 0; q < steps; ++q) acc += M.loss_of(hp.data() + (size_t)q * Mlp::kLossBlocks, batch);
    epoch_loss = acc / steps;
  }
  (void)hipStreamSynchronize(M.st);
  (void)hipFree(dPart);
  (void)hipFree(dX);
  (void)hipFree(dT);
  (void)hipFree(dIdx);
  if (last_loss) *last_loss = epoch_loss;
  return mlp_status(M);
}

// Predictions in the original output units (out_std * y + out_mean) for
// `rows` host input rows; Y is rows x nout.
int pdos_mlp_predict(PdosMlp* P, const float* X, int rows, float* Y, double* kernel_ms) {
  Mlp& M = P->net;
  std::string err;
  const int chunk = 8192;
  if (!M.ensure_rows(std::min(rows, chunk), err)) { g_mlp_err = err; return -1; }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float total = 0.f;
  for (int s = 0; s < rows; s += chunk) {
    const int r = std::min(chunk, rows - s);
    (void)hipMemcpyAsync(M.Xf, X + (size_t)s * M.nin, (size_t)r * M.nin * 4, hipMemcpyHostToDevice, M.st);
    (void)hipEventRecord(e0, M.st);
    M.forward(r, true);
    (void)hipEventRecord(e1, M.st);
    (void)hipMemcpy2DAsync(Y + (size_t)s * M.nout, M.nout * 4, M.Y, M.pout * 4, M.nout * 4, r,
                           hipMemcpyDeviceToHost, M.st);
    (void)hipStreamSynchronize(M.st);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    total += ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (kernel_ms) *kernel_ms = total;
  return mlp_status(M);
}

// Device-to-device prediction: dX (device, rows x nin raw features) to dY
// (device, rows x ldy, original output units).  The timed warm-start path.
int pdos_mlp_predict_dev(PdosMlp* P, const float* dX, int rows, float* dY, int ldy) {
  Mlp& M = P->net;
  std::string err;
  const int chunk = 8192;
  if (!M.ensure_rows(std::min(rows, chunk), err)) { g_mlp_err = err; return -1; }
  for (int s = 0; s < rows; s += chunk) {
    const int r = std::min(chunk, rows - s);
    (void)hipMemcpyAsync(M.Xf, dX + (size_t)s * M.nin, (size_t)r * M.nin * 4, hipMemcpyDeviceToDevice, M.st);
    M.forward(r, true);
    (void)hipMemcpy2DAsync(dY + (size_t)s * ldy, (size_t)ldy * 4, M.Y, M.pout * 4, M.nout * 4, r,
                           hipMemcpyDeviceToDevice, M.st);
  }
  return mlp_status(M);
}

// pdos_mlp_predict_dev without any allocation: the activations live in the
// file-scope `scratch` of this GPU, rows are processed in chunks that fit it.
// Each output row depends only on its own input row, so the results equal
// pdos_mlp_predict_dev's bit for bit.
int pdos_mlp_predict_dev_noalloc(PdosMlp* P, const float* dX, int rows, float* dY, int ldy) {
  using pdos_amd::pad64;
  if (!P) { g_mlp_err = "null network"; return -1; }
  Mlp& M = P->net;
  if (rows < 0 || (rows > 0 && (!dX || !dY)) || ldy < M.nout) { g_mlp_err = "bad arguments"; return -1; }
  std::string err;
  unsigned char* base = scratch_of(M.device, err);
  if (!base) { g_mlp_err = err; return -1; }
  auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
  // bytes of one chunk of rp rows: xb, z (shared by both layers), h1, h2, y
  auto need = [&](size_t rp) {
    return al(rp * M.pin * 2) + al(rp * M.ph * 4) + 2 * al(rp * M.ph * 2) + al(rp * M.pout * 4);
  };
  if (need(64) > kMlpScratchBytes) {
    g_mlp_err = "network too wide for the no-alloc scratch: 64 rows need " + std::to_string(need(64)) +
                " bytes, the scratch holds " + std::to_string(kMlpScratchBytes);
    return -1;
  }
  int chunk = 64;
  while (need((size_t)chunk + 64) <= kMlpScratchBytes && chunk < (1 << 20)) chunk += 64;
  for (int s = 0; s < rows; s += chunk) {
    const int r = std::min(chunk, rows - s);
    const size_t rp = (size_t)pad64(r);
    __bf16* xb = reinterpret_cast<__bf16*>(base);
    float* z = reinterpret_cast<float*>(base + al(rp * M.pin * 2));
    __bf16* h1 = reinterpret_cast<__bf16*>(reinterpret_cast<unsigned char*>(z) + al(rp * M.ph * 4));
    __bf16* h2 = reinterpret_cast<__bf16*>(reinterpret_cast<unsigned char*>(h1) + al(rp * M.ph * 2));
    float* y = reinterpret_cast<float*>(reinterpret_cast<unsigned char*>(h2) + al(rp * M.ph * 2));
    M.forward_into(r, true, dX + (size_t)s * M.nin, xb, z, h1, z, h2, y);
    const hipError_t e = hipMemcpy2DAsync(dY + (size_t)s * ldy, (size_t)ldy * 4, y, M.pout * 4, M.nout * 4, r,
                                          hipMemcpyDeviceToDevice, M.st);
    if (e != hipSuccess) { g_mlp_err = std::string("copy: ") + hipGetErrorString(e); return -1; }
  }
  return mlp_status(M);
}

}  // extern "C"
