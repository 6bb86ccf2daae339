// Complex FP64 GEMM on CDNA matrix cores: C[MxN] = A[MxK] * B[KxN], row-major,
// complex128 interleaved.
//
// Each 16x16 complex output tile keeps a real and an imaginary accumulator
// and takes four v_mfma_f64_16x16x4_f64 per 4-deep k step:
//   Cr += Ar*Br, Cr += (-Ai)*Bi, Ci += Ar*Bi, Ci += Ai*Br.
// A workgroup of 4 waves owns a 64x64 complex tile (each wave 32x32 = 2x2
// MFMA tiles); operands are staged through LDS 16 complex deep.
#include <algorithm>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int kThreads = 256;
constexpr int LDA_S = BK + 1;  // padded LDS row of the A tile (complex elements)
constexpr int LDB_S = BN + 1;  // padded LDS row of the B tile

__global__ void __launch_bounds__(kThreads) zgemm_f64_kernel(const c128 *__restrict__ A, const c128 *__restrict__ B,
                                                             c128 *__restrict__ C, uint64_t M, uint64_t N, uint64_t K,
                                                             uint64_t lda, uint64_t ldb, uint64_t ldc) {
  __shared__ c128 As[BM * LDA_S];
  __shared__ c128 Bs[BK * LDB_S];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32;  // wave tile origin inside the block tile
  const int wn = (wave & 1) * 32;
  const uint64_t ntn = (N + BN - 1) / BN;
  const uint64_t m0 = ((uint64_t)blockIdx.x / ntn) * BM;
  const uint64_t n0 = ((uint64_t)blockIdx.x % ntn) * BN;

  double4_t accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
    }

  const c128 zero = {0.0, 0.0};
  for (uint64_t k0 = 0; k0 < K; k0 += BK) {
    // Stage A (64 x 16) and B (16 x 64): 4 complex elements per thread each.
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int idx = tid + r * kThreads;
      int ar = idx / BK, ac = idx % BK;
      uint64_t gm = m0 + ar, gk = k0 + ac;
      As[ar * LDA_S + ac] = (gm < M && gk < K) ? A[gm * lda + gk] : zero;
      int br = idx / BN, bc = idx % BN;
      uint64_t gk2 = k0 + br, gn = n0 + bc;
      Bs[br * LDB_S + bc] = (gk2 < K && gn < N) ? B[gk2 * ldb + gn] : zero;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4);
      c128 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[(wm + i * 16 + (lane & 15)) * LDA_S + kl];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kl * LDB_S + wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double nai = -a[i].im;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].re, b[j].re, accR[i][j], 0, 0, 0);
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b[j].im, accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].re, b[j].im, accI[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].im, b[j].re, accI[i][j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // Result layout of the 16x16 f64 MFMA (measured on gfx950): lane l holds
  // column l % 16 of rows l / 16 + 4 v, v = 0..3.
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        uint64_t gm = m0 + wm + i * 16 + (lane >> 4) + 4 * v;
        uint64_t gn = n0 + wn + j * 16 + (lane & 15);
        if (gm < M && gn < N) {
          c128 out;
          out.re = accR[i][j][v];
          out.im = accI[i][j][v];
          C[gm * ldc + gn] = out;
        }
      }
}

}  // namespace

qtnh_status launch_zgemm_f64(qtnh_ctx_t ctx, const c128 *A, const c128 *B, c128 *C, uint64_t M, uint64_t N, uint64_t K,
                             uint64_t lda, uint64_t ldb, uint64_t ldc) {
  const uint64_t tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  if (tiles > 0x7fffffffull) return set_error(ctx, QTNH_EUNSUPPORTED, "zgemm: grid too large");
  hipLaunchKernelGGL(zgemm_f64_kernel, dim3((unsigned)tiles), dim3(kThreads), 0, ctx->stream, A, B, C, M, N, K, lda,
                     ldb, ldc);
  return check_hip(ctx, hipGetLastError(), "zgemm launch");
}

}  // namespace qtnh_hip
