// Split-precision complex GEMM: error-free slicing onto INT8 matrix cores
// (the scheme of Ozaki, Ogita & Oishi, 2012, with integer slices as in
// Ootomo, Ozaki & Yokota, 2024), FP64 recombination.
//
// The complex product runs as one real GEMM on the embedded matrices
//   A_e (M x 2K) = A's natural interleaved layout (re, im per element),
//   B_e (2K x 2N) with 2x2 blocks [[Br, Bi], [-Bi, Br]],
// so that C_e (M x 2N) is C's natural layout.
//
// Each row of A_e (and each column pair of B_e) is scaled by a power of two,
// x = 2^e R with |R| < 1.  Y = rint(R 2^(8S-2)) is an integer of at most
// 8S-2 bits, written exactly as S balanced base-256 digits in [-128, 127]:
//   R = sum_p q_p 2^-(6 + 8p) + r,  |r| <= 2^-(8S-1).
// Products of digits are exact in int32 (v_mfma_i32_16x16x64_i8).  The
// kernel keeps one int32 accumulator per diagonal d = p + q < S (the
// triangular truncation) and recombines
//   C_ij = 2^(ea_i + eb_j) sum_d 2^-(12 + 8d) acc_d
// in FP64 at the end.  Relative error on Gaussian complex inputs: about
// 7e-11 with S = 5 (15 digit products), 2e-13 with S = 6 (21).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

typedef int int4_t __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kMaxSlices = 8;

// Slice planes are stored in the order the GEMM kernels keep operands in LDS:
// 1 KiB chunks of 16 rows x 64 bytes of depth, the K2 / 64 chunks of a row
// group in depth order, row groups in row order.  Inside a chunk, 16-byte
// slot 4 rr + (c ^ (rr / 2 % 4)) holds depth bytes 16 c .. 16 c + 15 of row
// rr (the XOR spreads the fragment reads of 16 rows over distinct LDS banks).
// A wave instruction then copies one whole chunk, 1 KiB contiguous, straight
// into LDS.  Byte offset of real element (r, k) in a plane of depth K2:
__device__ __forceinline__ uint64_t tiled_at(uint64_t r, uint64_t k, uint64_t K2) {
  const uint64_t rr = r & 15, c = (k >> 4) & 3;
  return ((((r >> 4) * (K2 >> 6)) + (k >> 6)) << 10) + ((rr * 4 + (c ^ ((rr >> 1) & 3))) << 4) + (k & 15);
}

// One block per complex row.  mode 0 (A): real row i = (re_0, im_0, re_1, ...).
// mode 1 (B transposed, row j of B^T): real rows 2j = (re, -im, ...), 2j+1 = (im, re, ...).
__global__ void __launch_bounds__(kThreads) split_rows_kernel(const c128 *__restrict__ X, uint64_t rows, uint64_t K,
                                                              uint64_t ldx, int mode, int S, int8_t *__restrict__ out,
                                                              uint64_t plane, uint64_t ldo, int *__restrict__ expo) {
  __shared__ double red[kThreads];
  const uint64_t row = blockIdx.x;
  if (row >= rows) return;  // uniform per block
  const c128 *x = X + row * ldx;
  double m = 0.0;
  for (uint64_t k = threadIdx.x; k < K; k += kThreads) {
    const c128 v = x[k];
    m = fmax(m, fmax(fabs(v.re), fabs(v.im)));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  int e = 0;
  if (red[0] > 0.0) frexp(red[0], &e);  // red[0] = f 2^e, f in [0.5, 1): |x| 2^-e < 1
  if (threadIdx.x == 0) {
    if (mode == 0) {
      expo[row] = e;
    } else {
      expo[2 * row] = e;
      expo[2 * row + 1] = e;
    }
  }
  // Balanced base-256 digits of rint(v 2^(8S-2-e)), most significant first.
  auto digits = [&](double v, int8_t *d) {
    long long y = (long long)rint(ldexp(v, 8 * S - 2 - e));
#pragma unroll
    for (int p = kMaxSlices - 1; p >= 0; --p) {
      if (p >= S) continue;
      const long long t = ((y + 128) & 255) - 128;
      d[p] = (int8_t)t;
      y = (y - t) >> 8;  // exact: y - t is a multiple of 256
    }
  };
  auto pack = [](int8_t lo, int8_t hi) { return (uint16_t)((uint8_t)lo | ((uint16_t)(uint8_t)hi << 8)); };
  for (uint64_t k = threadIdx.x; k < K; k += kThreads) {
    const c128 v = x[k];
    int8_t dr[kMaxSlices], di[kMaxSlices], dn[kMaxSlices];
    digits(v.re, dr);
    digits(v.im, di);
    if (mode == 1) digits(-v.im, dn);
#pragma unroll
    for (int p = 0; p < kMaxSlices; ++p) {
      if (p >= S) continue;
      int8_t *o = out + (uint64_t)p * plane;
      if (mode == 0) {
        *reinterpret_cast<uint16_t *>(o + tiled_at(row, 2 * k, ldo)) = pack(dr[p], di[p]);
      } else {
        *reinterpret_cast<uint16_t *>(o + tiled_at(2 * row, 2 * k, ldo)) = pack(dr[p], dn[p]);
        *reinterpret_cast<uint16_t *>(o + tiled_at(2 * row + 1, 2 * k, ldo)) = pack(di[p], dr[p]);
      }
    }
  }
}

// Block tile: 8 waves in a WGM x WGN grid, each owning WM x WN outputs
// (real rows x real columns): 128 x 128 with 64 x 32 wave tiles for S <= 5,
// 128 x 64 with 32 x 32 wave tiles above.  Operand panels (S slices x rows x
// 64 bytes of depth) go from global memory straight into LDS
// (global_load_lds_dwordx4: each wave instruction fills 1 KiB = 16 rows x 64
// bytes, the lanes choosing source addresses so that the 16-byte chunks land
// XOR-swizzled by row and the fragment reads hit distinct banks).
// Two LDS stages when they fit: the next depth step streams in while the
// current one feeds the matrix cores.
constexpr int BK = 64;
constexpr int kGroupM = 4;  // tile rows per group of the tile order
constexpr int kLdsBytes = 160 * 1024;

// LDS slot (16-byte units) of logical chunk c of row r of slice p.  Fragment
// reads take 16 rows at one logical chunk; LDS serves them 8 lanes (128
// bytes over 32 banks) at a time, so rows r and r + 2, r + 4, r + 6 must use
// different physical chunks: physical = logical ^ ((r / 2) % 4).
__device__ __forceinline__ int lds_at(int p, int r, int c, int rows) { return ((p * rows + r) << 2) + (c ^ ((r >> 1) & 3)); }

template <int S, int WM, int WN, int WGM, int WGN>
struct OzShape {
  static constexpr int BM = WM * WGM, BN = WN * WGN;
  static constexpr int FA = WM / 16, FB = WN / 16;
  static constexpr int NST = (2 * S * (BM + BN) * BK <= kLdsBytes) ? 2 : 1;
  static constexpr int CA = S * BM / 16, CB = S * BN / 16;  // 1 KiB chunks per stage
  static constexpr int NW = WGM * WGN, THREADS = 64 * NW;
};

// C (M x N2 real, leading dim ldc in doubles) (+)= recombined slice products.
// As: S tiled planes of M x K2 int8 (depth lda of the whole plane; the base
// already points at this launch's depth chunk); Bt: S tiled planes of N2 x K2.
template <int S, int WM, int WN, int WGM, int WGN, int MINW>
__global__ void __launch_bounds__(64 * WGM * WGN, MINW) ozaki_gemm_kernel(
    const int8_t *__restrict__ As, uint64_t a_plane, uint64_t lda, const int8_t *__restrict__ Bt, uint64_t b_plane,
    uint64_t ldb, const int *__restrict__ ea, const int *__restrict__ eb, double *__restrict__ C, uint64_t ldc,
    uint64_t M, uint64_t N2, uint64_t K2, unsigned tilesM, unsigned tilesN, int accumulate) {
  using Z = OzShape<S, WM, WN, WGM, WGN>;
  constexpr int BM = Z::BM, BN = Z::BN, NST = Z::NST;
  __shared__ int4_t sA[NST][S * BM * 4];
  __shared__ int4_t sB[NST][S * BN * 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;

  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
  // range of tiles, walked in groups of kGroupM tile rows, so the blocks an XCD
  // runs together share A and B panels in its L2.
  const unsigned T = gridDim.x, b = blockIdx.x;
  const unsigned xcd = b & 7, per = T >> 3, rem = T & 7;
  const unsigned t = xcd * per + min(xcd, rem) + (b >> 3);
  const unsigned gsize = kGroupM * tilesN;
  const unsigned first_m = (t / gsize) * kGroupM;
  const unsigned gm_cnt = min((unsigned)kGroupM, tilesM - first_m);
  const unsigned in = t % gsize;
  const uint64_t m0 = (uint64_t)(first_m + in % gm_cnt) * BM;
  const uint64_t n0 = (uint64_t)(in / gm_cnt) * BN;

  int4_t acc[S][Z::FA][Z::FB];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < Z::FA; ++i)
#pragma unroll
      for (int j = 0; j < Z::FB; ++j) acc[d][i][j] = int4_t{0, 0, 0, 0};

  // A chunk of the tiled planes is one wave instruction: lane i copies its
  // 16-byte slot i into LDS slot i.  The planes hold whole tiles of rows (the
  // rows past M or N2 are padding whose products land only in outputs that are
  // not stored), so no row is clamped: a chunk's address is a wave-uniform base
  // plus the lane's 16 i, and issuing a chunk costs no per-lane arithmetic.
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t off = 16 * lane;
  // Row groups of 16 rows are lda / 64 chunks of 1 KiB apart; depth step kt is
  // chunk kt of its row group.
  auto issue = [&](int st, uint64_t kt) {
#pragma unroll
    for (int i = 0; i < (Z::CA + Z::CB + Z::NW - 1) / Z::NW; ++i) {
      const int c = wv + Z::NW * i;
      if (c < Z::CA + Z::CB) {
        const bool isA = c < Z::CA;
        const int cc = isA ? c : c - Z::CA;
        const int rows = isA ? BM : BN;
        const int p = cc / (rows / 16), g = cc % (rows / 16);
        const int8_t *base = isA ? As + p * a_plane + ((m0 >> 4) + g) * (lda << 4) + (kt << 10)
                                 : Bt + p * b_plane + ((n0 >> 4) + g) * (ldb << 4) + (kt << 10);
        int4_t *dst = isA ? &sA[st][(p * rows + g * 16) * 4] : &sB[st][(p * rows + g * 16) * 4];
        __builtin_amdgcn_global_load_lds((const void *)(base + off), (void __attribute__((address_space(3))) *)dst, 16,
                                         0, 0);
      }
    }
  };

  // The caller pads the real depth K2 to a multiple of BK with zeros.
  const uint64_t nk = K2 / BK;
  issue(0, 0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (uint64_t kt = 0; kt < nk; ++kt) {
    const int cur = (NST == 2) ? (int)(kt & 1) : 0;
    if (NST == 2 && kt + 1 < nk) issue(cur ^ 1, kt + 1);  // streams in during the MFMAs below
    // Fragments (16x16x64 i8): lane l holds row / column l % 16, k bytes 16 (l / 16) .. +15.
    // All B slices stay in registers; A slices stream through, one at a time.
    int4_t fb[S][Z::FB];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int j = 0; j < Z::FB; ++j) fb[q][j] = sB[cur][lds_at(q, wn + 16 * j + (lane & 15), lane >> 4, BN)];
#pragma unroll
    for (int p = 0; p < S; ++p) {
      int4_t fa[Z::FA];
#pragma unroll
      for (int i = 0; i < Z::FA; ++i) fa[i] = sA[cur][lds_at(p, wm + 16 * i + (lane & 15), lane >> 4, BM)];
#pragma unroll
      for (int q = 0; q < S - p; ++q)
#pragma unroll
        for (int i = 0; i < Z::FA; ++i)
#pragma unroll
          for (int j = 0; j < Z::FB; ++j)
            acc[p + q][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[q][j], acc[p + q][i][j], 0, 0, 0);
    }
    if (NST == 1 && kt + 1 < nk) {
      __syncthreads();  // everyone is done reading the only stage
      issue(0, kt + 1);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
  // 16x16 i32 result layout: lane l holds column l % 16 of rows 4 (l / 16) + v.
#pragma unroll
  for (int i = 0; i < Z::FA; ++i)
#pragma unroll
    for (int j = 0; j < Z::FB; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint64_t gm = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
        const uint64_t gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N2) {
          double sum = 0.0;
#pragma unroll
          for (int d = S - 1; d >= 0; --d) sum += ldexp((double)acc[d][i][j][v], -(12 + 8 * d));
          sum = ldexp(sum, ea[gm] + eb[gn]);
          double *c = C + gm * ldc + gn;
          *c = accumulate ? *c + sum : sum;
        }
      }
}

// Register-staged variant for S >= 6 (128 x 64 tiles, 8 waves of 32 x 32):
// each thread loads its 16-byte chunks of the next depth step into registers
// before the MFMAs of the current one and writes them to the other LDS stage
// after them.  Same arithmetic and result layout as ozaki_gemm_kernel.
template <int S>
__global__ void __launch_bounds__(512) ozaki_gemm_staged(
    const int8_t *__restrict__ As, uint64_t a_plane, uint64_t lda, const int8_t *__restrict__ Bt, uint64_t b_plane,
    uint64_t ldb, const int *__restrict__ ea, const int *__restrict__ eb, double *__restrict__ C, uint64_t ldc,
    uint64_t M, uint64_t N2, uint64_t K2, unsigned tilesM, unsigned tilesN, int accumulate) {
  constexpr int BM = 128, BN = 64, kGemmThreads = 512;
  constexpr int NST = (S <= 6) ? 2 : 1;
  __shared__ int4_t sA[NST][S * BM * 4];
  __shared__ int4_t sB[NST][S * BN * 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
  // range of tiles, walked in groups of kGroupM tile rows, so the blocks an XCD
  // runs together share A and B panels in its L2.
  const unsigned T = gridDim.x, b = blockIdx.x;
  const unsigned xcd = b & 7, per = T >> 3, rem = T & 7;
  const unsigned t = xcd * per + min(xcd, rem) + (b >> 3);
  const unsigned gsize = kGroupM * tilesN;
  const unsigned first_m = (t / gsize) * kGroupM;
  const unsigned gm_cnt = min((unsigned)kGroupM, tilesM - first_m);
  const unsigned in = t % gsize;
  const uint64_t m0 = (uint64_t)(first_m + in % gm_cnt) * BM;
  const uint64_t n0 = (uint64_t)(in / gm_cnt) * BN;

  int4_t acc[S][2][2];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[d][i][j] = int4_t{0, 0, 0, 0};

  // Staging: A chunk (row tid / 4, 16-byte chunk tid % 4) of every slice; B
  // chunks tid + 512 i over the S x 64 x 4 chunks of the B panel.
  constexpr int BL = (S * BN * 4 + kGemmThreads - 1) / kGemmThreads;
  const int4_t zero = {0, 0, 0, 0};
  int4_t pa[S], pb[BL];
  const int ar = tid >> 2, ac = tid & 3;
  auto fetch = [&](uint64_t k0) {
    const uint64_t ga = m0 + ar;
#pragma unroll
    for (int p = 0; p < S; ++p)
      pa[p] = (ga < M) ? *reinterpret_cast<const int4_t *>(As + p * a_plane + tiled_at(ga, k0 + ac * 16, lda)) : zero;
#pragma unroll
    for (int i = 0; i < BL; ++i) {
      const int c = tid + kGemmThreads * i;
      const int p = c / (BN * 4), r = (c % (BN * 4)) >> 2, ch = c & 3;
      const uint64_t gb = n0 + r;
      pb[i] = (c < S * BN * 4 && gb < N2)
                  ? *reinterpret_cast<const int4_t *>(Bt + p * b_plane + tiled_at(gb, k0 + ch * 16, ldb))
                  : zero;
    }
  };
  auto stash = [&](int st) {
#pragma unroll
    for (int p = 0; p < S; ++p) sA[st][lds_at(p, ar, ac, BM)] = pa[p];
#pragma unroll
    for (int i = 0; i < BL; ++i) {
      const int c = tid + kGemmThreads * i;
      if (c < S * BN * 4) sB[st][lds_at(c / (BN * 4), (c % (BN * 4)) >> 2, c & 3, BN)] = pb[i];
    }
  };

  // The caller pads the real depth K2 to a multiple of BK with zeros.
  const uint64_t nk = K2 / BK;
  fetch(0);
  stash(0);
  __syncthreads();
  for (uint64_t kt = 0; kt < nk; ++kt) {
    const int cur = (NST == 2) ? (int)(kt & 1) : 0;
    if (kt + 1 < nk) fetch((kt + 1) * BK);  // in flight during the MFMAs below
    // Fragments (16x16x64 i8): lane l holds row / column l % 16, k bytes 16 (l / 16) .. +15.
    // All B slices stay in registers; A slices stream through, one at a time.
    int4_t fb[S][2];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int j = 0; j < 2; ++j) fb[q][j] = sB[cur][lds_at(q, wn + 16 * j + (lane & 15), lane >> 4, BN)];
#pragma unroll
    for (int p = 0; p < S; ++p) {
      int4_t fa[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) fa[i] = sA[cur][lds_at(p, wm + 16 * i + (lane & 15), lane >> 4, BM)];
#pragma unroll
      for (int q = 0; q < S - p; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[p + q][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[q][j], acc[p + q][i][j], 0, 0, 0);
    }
    if (NST == 2) {
      if (kt + 1 < nk) stash(cur ^ 1);
      __syncthreads();
    } else {
      __syncthreads();
      if (kt + 1 < nk) {
        stash(0);
        __syncthreads();
      }
    }
  }
  // 16x16 i32 result layout: lane l holds column l % 16 of rows 4 (l / 16) + v.
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint64_t gm = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
        const uint64_t gn = n0 + wn + 16 * j + (lane & 15);
        if (gm < M && gn < N2) {
          double sum = 0.0;
#pragma unroll
          for (int d = S - 1; d >= 0; --d) sum += ldexp((double)acc[d][i][j][v], -(12 + 8 * d));
          sum = ldexp(sum, ea[gm] + eb[gn]);
          double *c = C + gm * ldc + gn;
          *c = accumulate ? *c + sum : sum;
        }
      }
}


// Tile shapes of ozaki_gemm_kernel (wave tile WM x WN, WGM x WGN waves, MINW
// waves per SIMD).  0: 128 x 128 block tiles of 8 waves, one workgroup per
// compute unit (both LDS stages fill the 160 KiB); 1: 64 x 64 of 4 waves, two
// workgroups per compute unit (80 KiB each), so one's operand DMA overlaps the
// other's matrix-core work; 2: 128 x 64 of 8 waves (120 KiB); 3: 64 x 64 of 8
// waves with 32 x 16 wave tiles, two workgroups per compute unit.  Every shape
// computes the same exact integer sums, so results do not depend on it.
static int g_split_tile = 0;

template <int S, int WM, int WN, int WGM, int WGN, int MINW>
void launch_ozaki_shape(hipStream_t st, const int8_t *As, uint64_t ap, uint64_t lda, const int8_t *Bt, uint64_t bp,
                        uint64_t ldb, const int *ea, const int *eb, double *C, uint64_t ldc, uint64_t M, uint64_t N2,
                        uint64_t K2, int accumulate) {
  using Z = OzShape<S, WM, WN, WGM, WGN>;
  const unsigned tilesM = (unsigned)((M + Z::BM - 1) / Z::BM), tilesN = (unsigned)((N2 + Z::BN - 1) / Z::BN);
  hipLaunchKernelGGL((ozaki_gemm_kernel<S, WM, WN, WGM, WGN, MINW>), dim3(tilesM * tilesN), dim3(Z::THREADS), 0, st,
                     As, ap, lda, Bt, bp, ldb, ea, eb, C, ldc, M, N2, K2, tilesM, tilesN, accumulate);
}

template <int S>
void launch_ozaki(hipStream_t st, const int8_t *As, uint64_t ap, uint64_t lda, const int8_t *Bt, uint64_t bp,
                  uint64_t ldb, const int *ea, const int *eb, double *C, uint64_t ldc, uint64_t M, uint64_t N2,
                  uint64_t K2, int accumulate) {
  if constexpr (S >= 6) {
    const unsigned tilesM = (unsigned)((M + 127) / 128), tilesN = (unsigned)((N2 + 63) / 64);
    hipLaunchKernelGGL((ozaki_gemm_staged<S>), dim3(tilesM * tilesN), dim3(512), 0, st, As, ap, lda, Bt, bp, ldb, ea,
                       eb, C, ldc, M, N2, K2, tilesM, tilesN, accumulate);
  } else {
    switch (g_split_tile) {
      case 1:
        launch_ozaki_shape<S, 32, 32, 2, 2, 2>(st, As, ap, lda, Bt, bp, ldb, ea, eb, C, ldc, M, N2, K2, accumulate);
        break;
      case 2:
        launch_ozaki_shape<S, 64, 32, 2, 2, 1>(st, As, ap, lda, Bt, bp, ldb, ea, eb, C, ldc, M, N2, K2, accumulate);
        break;
      case 3:
        launch_ozaki_shape<S, 32, 16, 2, 4, 4>(st, As, ap, lda, Bt, bp, ldb, ea, eb, C, ldc, M, N2, K2, accumulate);
        break;
      default:
        launch_ozaki_shape<S, 64, 32, 2, 4, 1>(st, As, ap, lda, Bt, bp, ldb, ea, eb, C, ldc, M, N2, K2, accumulate);
        break;
    }
  }
}

}  // namespace

// Number of 8-bit slices of the split-precision GEMM (default 5).
static int g_split_slices = 5;

// Rows of a slice plane: whole 128-row tiles (every block tile is at most 128
// rows of A and 128 rows of B^T), so the GEMM kernels read padding rows past
// the end instead of clamping row indices.  The padding rows are never
// written; their products reach only outputs past M or N2, which are not stored.
static uint64_t plane_rows(uint64_t rows) { return (rows + 127) / 128 * 128; }

extern "C" void qtnh_tune_split(int slices) { g_split_slices = std::max(1, std::min(slices, kMaxSlices)); }

extern "C" void qtnh_tune_split_tile(int shape) { g_split_tile = (shape >= 0 && shape <= 3) ? shape : 0; }

int split_slices() { return g_split_slices; }

// A-priori componentwise bound of the truncated scheme: each scaled value is
// within 2^-(8S-1) of its digits, the dropped diagonals d >= S add at most
// (S + 1) 2^(2-8S) per real product, so with 2^e <= 2 max|x| for both scales
//   |C_ij - C~_ij| <= eta max_k |A_ik| max_k |B_kj|,  eta = 16 (S + 2) 2^-8S (2K).
double split_error_bound(int S, uint64_t K) { return 16.0 * (double)(S + 2) * ldexp(1.0, -8 * S) * (double)(2 * K); }

qtnh_status launch_zgemm_split(qtnh_ctx_t ctx, const c128 *A, const c128 *Bt, c128 *C, uint64_t M, uint64_t N,
                               uint64_t K, void *scratch, uint64_t scratch_bytes, double *err_out) {
  const int S = g_split_slices;
  const uint64_t K2 = ((2 * K + 63) / 64) * 64;  // padded real depth (zeros beyond 2K)
  const uint64_t N2 = 2 * N;
  const uint64_t a_plane = plane_rows(M) * K2, b_plane = plane_rows(N2) * K2;
  const uint64_t need = S * (a_plane + b_plane) + 4 * (M + N2) + 512;
  if (scratch_bytes < need) return set_error(ctx, QTNH_EINVAL, "split: scratch too small");
  if (M > 0x7fffffffull || N > 0x7fffffffull) return set_error(ctx, QTNH_EUNSUPPORTED, "split: too many rows");
  int8_t *As = (int8_t *)scratch;
  int8_t *Bs = As + S * a_plane;
  int *ea = (int *)(((uintptr_t)(Bs + S * b_plane) + 255) & ~(uintptr_t)255);
  int *eb = ea + M;
  // Padding columns of the slice planes must read as zero.
  if (K2 != 2 * K) {
    qtnh_status st = launch_zero(ctx, As, S * (a_plane + b_plane));
    if (st) return st;
  }
  hipLaunchKernelGGL(split_rows_kernel, dim3((unsigned)M), dim3(kThreads), 0, ctx->stream, A, M, K, K, 0, S, As, a_plane,
                     K2, ea);
  qtnh_status st = check_hip(ctx, hipGetLastError(), "split A launch");
  if (st) return st;
  hipLaunchKernelGGL(split_rows_kernel, dim3((unsigned)N), dim3(kThreads), 0, ctx->stream, Bt, N, K, K, 1, S, Bs,
                     b_plane, K2, eb);
  st = check_hip(ctx, hipGetLastError(), "split B launch");
  if (st) return st;
  // int32 accumulators stay exact while (S pairs) * 128 * 128 * depth < 2^31:
  // depth chunks of 2^13 (S <= 8).
  const uint64_t chunk = 8192;
  if (((M + 127) / 128) * ((N2 + 63) / 64) > 0x7fffffffull)
    return set_error(ctx, QTNH_EUNSUPPORTED, "split: grid too large");
  // Depth chunk k0 (a multiple of 64) starts at chunk k0 / 64 of every row
  // group of the tiled planes: byte 16 k0.
  for (uint64_t k0 = 0; k0 < K2; k0 += chunk) {
    const uint64_t kc = std::min(chunk, K2 - k0);
    const int accumulate = k0 > 0;
    switch (S) {
#define QTNH_SPLIT_CASE(s)                                                                                            \
  case s:                                                                                                            \
    launch_ozaki<s>(ctx->stream, As + 16 * k0, a_plane, K2, Bs + 16 * k0, b_plane, K2, ea, eb, (double *)C, N2, M,   \
                    N2, kc, accumulate);                                                                             \
    break;
      QTNH_SPLIT_CASE(1)
      QTNH_SPLIT_CASE(2)
      QTNH_SPLIT_CASE(3)
      QTNH_SPLIT_CASE(4)
      QTNH_SPLIT_CASE(5)
      QTNH_SPLIT_CASE(6)
      QTNH_SPLIT_CASE(7)
      default:
        launch_ozaki<8>(ctx->stream, As + 16 * k0, a_plane, K2, Bs + 16 * k0, b_plane, K2, ea, eb, (double *)C, N2, M,
                        N2, kc, accumulate);
        break;
#undef QTNH_SPLIT_CASE
    }
    st = check_hip(ctx, hipGetLastError(), "split gemm launch");
    if (st) return st;
  }
  if (err_out) *err_out = split_error_bound(S, K);
  return QTNH_OK;
}

uint64_t split_scratch_bytes(uint64_t M, uint64_t N, uint64_t K) {
  const int S = g_split_slices;
  const uint64_t K2 = ((2 * K + 63) / 64) * 64;
  return S * (plane_rows(M) * K2 + plane_rows(2 * N) * K2) + 4 * (M + 2 * N) + 512;
}

}  // namespace qtnh_hip
