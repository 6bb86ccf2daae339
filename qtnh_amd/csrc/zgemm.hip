// Complex FP64 GEMM on CDNA matrix cores: C[MxN] = A[MxK] * B[KxN], row-major,
// complex128 interleaved.
//
// Each 16x16 complex output tile keeps a real and an imaginary accumulator
// and takes four v_mfma_f64_16x16x4_f64 per 4-deep k step:
//   Cr += Ar*Br, Cr += (-Ai)*Bi, Ci += Ar*Bi, Ci += Ai*Br.
// A workgroup of 4 waves owns a BM x BN complex tile, each wave a 32 x (BN/2)
// sub-tile (2 x BN/32 MFMA tiles).  Operands are staged through LDS BK complex
// deep; the global loads of stage s+1 are issued into registers before the
// MFMAs of stage s, so memory latency hides behind the (slow, 64-cycle) f64
// MFMAs.
#include <algorithm>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef double double2_t __attribute__((ext_vector_type(2)));

constexpr int BK = 16;
constexpr int kThreads = 256;

template <int BM, int BN>
struct ZTile {
  static constexpr int LDA = BK + 1;  // padded LDS rows (complex elements)
  static constexpr int LDB = BN + 1;
  static constexpr int A_PER = BM * BK / kThreads;  // complex loads per thread per stage
  static constexpr int B_PER = BK * BN / kThreads;
  static constexpr int WN = BN / 2;                 // wave tile: 32 x WN
  static constexpr int TN = WN / 16;                // MFMA tiles along n per wave
  static_assert(BM == 64, "two waves of 32 rows");
  static_assert(A_PER >= 1 && B_PER >= 1, "tile too small");
};

// One BM x BN output tile `tile` (row-major over the tile grid) of
// C = A * B over depth chunk `chunk` (k in [chunk kc, (chunk + 1) kc)), written
// to C + chunk cstride.  As / Bs: the workgroup's LDS staging buffers.
template <int BM, int BN>
__device__ __forceinline__ void zgemm_tile(const c128 *__restrict__ A, const c128 *__restrict__ B,
                                           c128 *__restrict__ C, uint64_t M, uint64_t N, uint64_t K, uint64_t lda,
                                           uint64_t ldb, uint64_t ldc, uint64_t kc, uint64_t cstride, uint64_t tile,
                                           uint64_t chunk, double2_t *As, double2_t *Bs) {
  using Z = ZTile<BM, BN>;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32;  // wave tile origin inside the block tile
  const int wn = (wave & 1) * Z::WN;
  const uint64_t ntn = (N + BN - 1) / BN;
  const uint64_t m0 = (tile / ntn) * BM;
  const uint64_t n0 = (tile % ntn) * BN;
  const uint64_t kb = chunk * kc;
  const uint64_t ke = kb + kc < K ? kb + kc : K;
  C += chunk * cstride;

  double4_t accR[2][Z::TN], accI[2][Z::TN];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < Z::TN; ++j) {
      accR[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};
    }

  const double2_t zero = {0.0, 0.0};
  const double2_t *A2 = reinterpret_cast<const double2_t *>(A);
  const double2_t *B2 = reinterpret_cast<const double2_t *>(B);
  double2_t pa[Z::A_PER], pb[Z::B_PER];
  auto fetch = [&](uint64_t k0) {
#pragma unroll
    for (int r = 0; r < Z::A_PER; ++r) {
      const int idx = tid + r * kThreads;
      const int ar = idx / BK, ac = idx % BK;
      const uint64_t gm = m0 + ar, gk = k0 + ac;
      pa[r] = (gm < M && gk < ke) ? A2[gm * lda + gk] : zero;
    }
#pragma unroll
    for (int r = 0; r < Z::B_PER; ++r) {
      const int idx = tid + r * kThreads;
      const int br = idx / BN, bc = idx % BN;
      const uint64_t gk = k0 + br, gn = n0 + bc;
      pb[r] = (gk < ke && gn < N) ? B2[gk * ldb + gn] : zero;
    }
  };
  fetch(kb);
  for (uint64_t k0 = kb; k0 < ke; k0 += BK) {
#pragma unroll
    for (int r = 0; r < Z::A_PER; ++r) {
      const int idx = tid + r * kThreads;
      As[(idx / BK) * Z::LDA + idx % BK] = pa[r];
    }
#pragma unroll
    for (int r = 0; r < Z::B_PER; ++r) {
      const int idx = tid + r * kThreads;
      Bs[(idx / BN) * Z::LDB + idx % BN] = pb[r];
    }
    __syncthreads();
    if (k0 + BK < ke) fetch(k0 + BK);  // in flight during the MFMAs below
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4);
      double2_t a[2], b[Z::TN];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[(wm + i * 16 + (lane & 15)) * Z::LDA + kl];
#pragma unroll
      for (int j = 0; j < Z::TN; ++j) b[j] = Bs[kl * Z::LDB + wn + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double nai = -a[i].y;
#pragma unroll
        for (int j = 0; j < Z::TN; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].x, accR[i][j], 0, 0, 0);
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, b[j].y, accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b[j].y, accI[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b[j].x, accI[i][j], 0, 0, 0);
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
    for (int j = 0; j < Z::TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint64_t gm = m0 + wm + i * 16 + (lane >> 4) + 4 * v;
        const uint64_t gn = n0 + wn + j * 16 + (lane & 15);
        if (gm < M && gn < N) {
          c128 out;
          out.re = accR[i][j][v];
          out.im = accI[i][j][v];
          C[gm * ldc + gn] = out;
        }
      }
}

// blockIdx.x: output tile; blockIdx.y: depth chunk (one chunk: the whole
// depth, into C; several: partial tiles at C + chunk cstride).
template <int BM, int BN>
__global__ void __launch_bounds__(kThreads) zgemm_f64_kernel(const c128 *__restrict__ A, const c128 *__restrict__ B,
                                                             c128 *__restrict__ C, uint64_t M, uint64_t N, uint64_t K,
                                                             uint64_t lda, uint64_t ldb, uint64_t ldc, uint64_t kc,
                                                             uint64_t cstride) {
  using Z = ZTile<BM, BN>;
  __shared__ double2_t As[BM * Z::LDA];
  __shared__ double2_t Bs[BK * Z::LDB];
  zgemm_tile<BM, BN>(A, B, C, M, N, K, lda, ldb, ldc, kc, cstride, blockIdx.x, blockIdx.y, As, Bs);
}

// Index of the descriptor whose block range holds block blk (descriptors
// sorted by first block; the first one starts at 0).
template <typename D>
__device__ __forceinline__ int find_desc(const D *__restrict__ d, int nd, uint64_t blk) {
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].first <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Many GEMMs in one launch: each block of a grid-stride walk over every
// GEMM's (chunk, tile) blocks computes one tile, exactly as zgemm_f64_kernel.
template <int BN>
__global__ void __launch_bounds__(kThreads) zgemm_grouped_kernel(const ZGroupDesc *__restrict__ d, int nd,
                                                                 uint64_t total) {
  using Z = ZTile<64, BN>;
  __shared__ double2_t As[64 * Z::LDA];
  __shared__ double2_t Bs[BK * Z::LDB];
  for (uint64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    const ZGroupDesc &D = d[find_desc(d, nd, blk)];
    const uint64_t local = blk - D.first;
    zgemm_tile<64, BN>(D.A, D.B, D.C, D.M, D.N, D.K, D.K, D.N, D.N, D.kc, D.cstride, local % D.tiles,
                       local / D.tiles, As, Bs);
  }
}

// Sums of depth-chunk partials of many GEMMs: C = sum over chunks of part[c], in chunk order.
__global__ void zgemm_grouped_sum_kernel(const ZSumDesc *__restrict__ d, int nd, uint64_t total) {
  for (uint64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    const ZSumDesc &D = d[find_desc(d, nd, blk)];
    const uint64_t i = (blk - D.first) * blockDim.x + threadIdx.x;
    if (i >= D.MN) continue;
    const double2_t *part = (const double2_t *)D.part;
    double2_t s = part[i];
    for (uint64_t c = 1; c < D.chunks; ++c) {
      const double2_t v = part[c * D.MN + i];
      s.x = s.x + v.x;
      s.y = s.y + v.y;
    }
    ((double2_t *)D.C)[i] = s;
  }
}

constexpr int kSkThreads = 256;

// Partial sum s of output mn = m N + n: ordered per-thread sums over the
// chunk's terms, then a fixed tree in LDS; red: the workgroup's LDS buffer.
__device__ __forceinline__ void splitk_partial(const double2_t *__restrict__ A, const double2_t *__restrict__ Bt,
                                               uint64_t N, uint64_t K, uint64_t kc, uint64_t mn, uint64_t s,
                                               double2_t *__restrict__ out, double2_t *red) {
  const uint64_t m = mn / N, n = mn % N;
  const double2_t *a = A + m * K, *b = Bt + n * K;
  const uint64_t k1 = std::min(K, (s + 1) * kc);
  double re = 0.0, im = 0.0;
  for (uint64_t k = s * kc + threadIdx.x; k < k1; k += kSkThreads) {
    const double2_t x = a[k], y = b[k];
    double pr, pi;
    cmul_parts(x.x, x.y, y.x, y.y, &pr, &pi);
    re = re + pr;
    im = im + pi;
  }
  red[threadIdx.x] = double2_t{re, im};
  __syncthreads();
  for (int w = kSkThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double2_t u = red[threadIdx.x], v = red[threadIdx.x + w];
      red[threadIdx.x] = double2_t{u.x + v.x, u.y + v.y};
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
  __syncthreads();  // red is reused by the workgroup's next partial
}

// Block (output mn, chunk s).
__global__ void __launch_bounds__(kSkThreads) splitk_partial_kernel(const double2_t *__restrict__ A,
                                                                    const double2_t *__restrict__ Bt, uint64_t N,
                                                                    uint64_t K, uint64_t kc, uint64_t S,
                                                                    double2_t *__restrict__ part) {
  __shared__ double2_t red[kSkThreads];
  splitk_partial(A, Bt, N, K, kc, blockIdx.x / S, blockIdx.x % S, part + blockIdx.x, red);
}

__device__ __forceinline__ double2_t splitk_sum(const double2_t *__restrict__ part, uint64_t mn, uint64_t S) {
  double re = 0.0, im = 0.0;
  for (uint64_t s = 0; s < S; ++s) {
    const double2_t v = part[mn * S + s];
    re = re + v.x;
    im = im + v.y;
  }
  return double2_t{re, im};
}

__global__ void splitk_final_kernel(const double2_t *__restrict__ part, uint64_t MN, uint64_t S,
                                    double2_t *__restrict__ C) {
  for (uint64_t mn = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; mn < MN; mn += (uint64_t)gridDim.x * blockDim.x)
    C[mn] = splitk_sum(part, mn, S);
}

// Split-K partials of many contractions: blocks [first, first + MN S) of a
// descriptor are its (output, chunk) pairs.
__global__ void __launch_bounds__(kSkThreads) splitk_grouped_partial_kernel(const SkDesc *__restrict__ d, int nd,
                                                                            uint64_t total) {
  __shared__ double2_t red[kSkThreads];
  for (uint64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    const SkDesc &D = d[find_desc(d, nd, blk)];
    const uint64_t local = blk - D.first;
    splitk_partial((const double2_t *)D.A, (const double2_t *)D.Bt, D.N, D.K, D.kc, local / D.S, local % D.S,
                   (double2_t *)D.part + local, red);
  }
}

// Their final sums: element i of descriptor j in block (i / 256) + D.first2.
__global__ void splitk_grouped_final_kernel(const SkDesc *__restrict__ d, int nd, uint64_t total) {
  for (uint64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (d[mid].first2 <= blk)
        lo = mid;
      else
        hi = mid - 1;
    }
    const SkDesc &D = d[lo];
    const uint64_t mn = (blk - D.first2) * blockDim.x + threadIdx.x;
    if (mn < D.MN) ((double2_t *)D.C)[mn] = splitk_sum((const double2_t *)D.part, mn, D.S);
  }
}
}  // namespace

qtnh_status launch_zgemm_splitk(qtnh_ctx_t ctx, const c128 *A, const c128 *Bt, c128 *C, uint64_t M, uint64_t N,
                                uint64_t K, uint64_t chunks, c128 *partials) {
  const uint64_t MN = M * N;
  const uint64_t kc = splitk_chunk_len(K, chunks);
  if (MN * chunks > 0x7fffffffull) return set_error(ctx, QTNH_EUNSUPPORTED, "splitk: grid too large");
  hipLaunchKernelGGL(splitk_partial_kernel, dim3((unsigned)(MN * chunks)), dim3(kSkThreads), 0, ctx->stream,
                     (const double2_t *)A, (const double2_t *)Bt, N, K, kc, chunks, (double2_t *)partials);
  qtnh_status st = check_hip(ctx, hipGetLastError(), "splitk launch");
  if (st) return st;
  hipLaunchKernelGGL(splitk_final_kernel, dim3(grid_for(MN, 256, 4096)), dim3(256), 0, ctx->stream,
                     (const double2_t *)partials, MN, chunks, (double2_t *)C);
  return check_hip(ctx, hipGetLastError(), "splitk final launch");
}

// Block width: 0 picks by size (64 x 128 tiles run at occupancy 1 and win
// while 64 x 64 tiles would number under 2048, about 8 per compute unit;
// past that the occupancy-3 64 x 64 kernel hides latency better), or 64 / 128.
static int g_zgemm_bn = 0;

extern "C" void qtnh_tune_zgemm(int bn) { g_zgemm_bn = (bn == 64 || bn == 128) ? bn : 0; }

// Depth chunks of a GEMM with few output tiles: enough chunks that tiles x
// chunks reaches about kZgemmTargetBlocks workgroups, each summing at least
// kZgemmMinChunk terms.  Depends on the shape only (the network planner sizes
// workspaces with it), so every caller of a shape runs the same chunks.
constexpr uint64_t kZgemmTargetBlocks = 1024, kZgemmMinChunk = 256;

static int zgemm_bn(uint64_t M, uint64_t N) {
  int bn = g_zgemm_bn;
  if (bn == 0) bn = (((M + 63) / 64) * ((N + 63) / 64) < 2048) ? 128 : 64;
  if (N < 128) bn = 64;
  return bn;
}

uint64_t zgemm_chunks(uint64_t M, uint64_t N, uint64_t K) {
  const int bn = zgemm_bn(M, N);
  const uint64_t tiles = ((N + bn - 1) / bn) * ((M + 63) / 64);
  if (tiles == 0 || tiles * 2 > kZgemmTargetBlocks || K < 2 * kZgemmMinChunk) return 1;
  const uint64_t want = (kZgemmTargetBlocks + tiles - 1) / tiles;
  return std::max<uint64_t>(1, std::min<uint64_t>(want, K / kZgemmMinChunk));
}

// C = sum over chunks of part[c] (M x N each), in chunk order.
__global__ void zgemm_chunk_sum_kernel(const double2_t *__restrict__ part, double2_t *__restrict__ C, uint64_t M,
                                       uint64_t N, uint64_t ldc, uint64_t chunks) {
  const uint64_t MN = M * N;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < MN; i += (uint64_t)gridDim.x * blockDim.x) {
    double2_t s = part[i];
    for (uint64_t c = 1; c < chunks; ++c) {
      const double2_t v = part[c * MN + i];
      s.x = s.x + v.x;
      s.y = s.y + v.y;
    }
    C[(i / N) * ldc + i % N] = s;
  }
}

qtnh_status launch_zgemm_f64(qtnh_ctx_t ctx, const c128 *A, const c128 *B, c128 *C, uint64_t M, uint64_t N, uint64_t K,
                             uint64_t lda, uint64_t ldb, uint64_t ldc, c128 *partials) {
  const int bn = zgemm_bn(M, N);
  const uint64_t tiles = ((N + bn - 1) / bn) * ((M + 63) / 64);
  if (tiles > 0x7fffffffull) return set_error(ctx, QTNH_EUNSUPPORTED, "zgemm: grid too large");
  const uint64_t chunks = zgemm_chunks(M, N, K);
  if (chunks > 1 && !partials) return set_error(ctx, QTNH_EINVAL, "zgemm: depth chunks need a partial buffer");
  const uint64_t kc = chunks > 1 ? (K + chunks - 1) / chunks : K;
  const uint64_t kcr = chunks > 1 ? (kc + BK - 1) / BK * BK : K;  // chunk bounds on BK steps
  const uint64_t nchunks = chunks > 1 ? (K + kcr - 1) / kcr : 1;
  c128 *out = chunks > 1 ? partials : C;
  const uint64_t out_ld = chunks > 1 ? N : ldc;
  const uint64_t cstride = chunks > 1 ? M * N : 0;
  const dim3 grid((unsigned)tiles, (unsigned)nchunks);
  if (bn == 128)
    hipLaunchKernelGGL((zgemm_f64_kernel<64, 128>), grid, dim3(kThreads), 0, ctx->stream, A, B, out, M, N, K, lda, ldb,
                       out_ld, kcr, cstride);
  else
    hipLaunchKernelGGL((zgemm_f64_kernel<64, 64>), grid, dim3(kThreads), 0, ctx->stream, A, B, out, M, N, K, lda, ldb,
                       out_ld, kcr, cstride);
  qtnh_status st = check_hip(ctx, hipGetLastError(), "zgemm launch");
  if (st || nchunks <= 1) return st;
  hipLaunchKernelGGL(zgemm_chunk_sum_kernel, dim3(grid_for(M * N, 256, 65536)), dim3(256), 0, ctx->stream,
                     (const double2_t *)partials, (double2_t *)C, M, N, ldc, nchunks);
  return check_hip(ctx, hipGetLastError(), "zgemm chunk sum launch");
}


uint64_t splitk_chunk_len(uint64_t K, uint64_t chunks) {
  return ((K + chunks - 1) / chunks + kSkThreads - 1) / kSkThreads * kSkThreads;
}

int zgemm_block_n(uint64_t M, uint64_t N) { return zgemm_bn(M, N); }

// Depth chunks of launch_zgemm_f64 on BK steps: (chunk length, chunk count).
void zgemm_chunk_shape(uint64_t M, uint64_t N, uint64_t K, uint64_t *kc_out, uint64_t *n_out) {
  const uint64_t chunks = zgemm_chunks(M, N, K);
  const uint64_t kc = chunks > 1 ? (K + chunks - 1) / chunks : K;
  const uint64_t kcr = chunks > 1 ? (kc + BK - 1) / BK * BK : K;
  *kc_out = kcr;
  *n_out = chunks > 1 ? (K + kcr - 1) / kcr : 1;
}

qtnh_status launch_zgemm_grouped(qtnh_ctx_t ctx, int bn, const ZGroupDesc *d, int nd, uint64_t total) {
  if (nd <= 0 || total == 0) return QTNH_OK;
  const unsigned grid = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus * 8);
  if (bn == 128)
    hipLaunchKernelGGL(zgemm_grouped_kernel<128>, dim3(grid), dim3(kThreads), 0, ctx->stream, d, nd, total);
  else
    hipLaunchKernelGGL(zgemm_grouped_kernel<64>, dim3(grid), dim3(kThreads), 0, ctx->stream, d, nd, total);
  return check_hip(ctx, hipGetLastError(), "grouped zgemm launch");
}

qtnh_status launch_zgemm_grouped_sum(qtnh_ctx_t ctx, const ZSumDesc *d, int nd, uint64_t total) {
  if (nd <= 0 || total == 0) return QTNH_OK;
  const unsigned grid = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(zgemm_grouped_sum_kernel, dim3(grid), dim3(256), 0, ctx->stream, d, nd, total);
  return check_hip(ctx, hipGetLastError(), "grouped chunk sum launch");
}

qtnh_status launch_splitk_grouped(qtnh_ctx_t ctx, const SkDesc *d, int nd, uint64_t total_partial,
                                  uint64_t total_final) {
  if (nd <= 0) return QTNH_OK;
  unsigned grid = (unsigned)std::min<uint64_t>(total_partial, (uint64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(splitk_grouped_partial_kernel, dim3(grid), dim3(kSkThreads), 0, ctx->stream, d, nd,
                     total_partial);
  qtnh_status st = check_hip(ctx, hipGetLastError(), "grouped splitk launch");
  if (st) return st;
  grid = (unsigned)std::min<uint64_t>(total_final, (uint64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(splitk_grouped_final_kernel, dim3(grid), dim3(256), 0, ctx->stream, d, nd, total_final);
  return check_hip(ctx, hipGetLastError(), "grouped splitk final launch");
}

}  // namespace qtnh_hip
