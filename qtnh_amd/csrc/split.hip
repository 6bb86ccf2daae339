// Split-precision complex GEMM: error-free slicing onto INT8 matrix cores
// (the scheme of Ozaki, Ogita & Oishi, 2012, with integer slices as in
// Ootomo, Ozaki & Yokota, 2024), FP64 recombination.
//
// The complex product takes three real products (Gauss's form):
//   T1 = Ar Br,  T2 = Ai Bi,  T3 = (Ar + Ai)(Br + Bi),
//   Re C = T1 - T2,  Im C = (T3 - T1) - T2,
// 3/4 of the matrix-core work of the four real products.
//
// Each real operand X is sliced per row (A) or per column (B): with m the
// row's largest |x| = fm 2^e (fm in [0.5, 1)),
//   Y = rint(x 2^-e (c / fm)),  c = 127 * 256^(S-1),
// so |Y| <= c, written exactly as S balanced base-256 digits in [-128, 127]
// whose leading digit stays in [-127, 127] (the row maximum maps to 127 *
// 256^(S-1): the full int8 range, against a power-of-two scale that leaves
// the leading digit at half to a quarter of it; about 7x less error).  Then
//   x = 2^e (fm / c) (Y + rho),  |rho| <= 1/2.
// Products of digits are exact in int32 (v_mfma_i32_16x16x64_i8).  The
// kernel keeps one int32 accumulator per diagonal d = p + q < S (the
// triangular truncation) and recombines in FP64:
//   T_ij = 2^(ea_i + eb_j) (fa_i fb_j) sum_d 256^(2S-2-d) acc_d.
// Relative error on Gaussian complex inputs: about 1.2e-11 with S = 5 (3 x 15
// digit products) at 2048^3.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

typedef int int4_t __attribute__((ext_vector_type(4)));
typedef double double2_t __attribute__((ext_vector_type(2)));

constexpr int kThreads = 256;
constexpr int kMaxSlices = 8;
constexpr int kParts = 3;  // T1, T2, T3

// Slice planes are stored in the order the GEMM kernels keep operands in LDS:
// 1 KiB chunks of 16 rows x 64 bytes of depth, the Kp / 64 chunks of a row
// group in depth order, row groups in row order.  Inside a chunk, 16-byte
// slot 4 rr + (c ^ (rr / 2 % 4)) holds depth bytes 16 c .. 16 c + 15 of row
// rr (the XOR spreads the fragment reads of 16 rows over distinct LDS banks).
// A wave instruction then copies one whole chunk, 1 KiB contiguous, straight
// into LDS.  Byte offset of element (r, k) in a plane of depth Kp:
__device__ __forceinline__ uint64_t tiled_at(uint64_t r, uint64_t k, uint64_t Kp) {
  const uint64_t rr = r & 15, c = (k >> 4) & 3;
  return ((((r >> 4) * (Kp >> 6)) + (k >> 6)) << 10) + ((rr * 4 + (c ^ ((rr >> 1) & 3))) << 4) + (k & 15);
}

// Scale of a row (or column) whose largest |x| is m: x 2^-e in (-1, 1) and
// the factor f = c / fm, so that rint(x 2^-e f) stays within +-c.  A zero
// row gets e = 0, f = 0 (all digits 0).  The recombination factor is fm / c.
template <int S>
struct RowScale {
  int e;
  double f, back;
  __device__ __forceinline__ explicit RowScale(double m) {
    const double c = ldexp(127.0, 8 * (S - 1));
    e = 0;
    if (m > 0.0) {
      const double fm = frexp(m, &e);
      f = c / fm;
      back = fm / c;
    } else {
      f = 0.0;
      back = 0.0;
    }
  }
};

// Balanced base-256 digits of Y = rint(x 2^-e f), most significant first:
// digit p in byte p of the result.  Y is an integer-valued double with
// |Y| <= 127 * 256^(S-1); for S <= 7 (|Y| < 2^55, ulp at most 4) the steps
// q = floor((y + 128) / 256), t = y - 256 q are exact in binary64, so the
// digits equal those of the int64 recurrence used for S = 8.
template <int S>
__device__ __forceinline__ uint64_t split_digits(double x, const RowScale<S> &sc) {
  uint64_t pk = 0;
  const double yd = rint(ldexp(x, -sc.e) * sc.f);
  if constexpr (S <= 7) {
    double y = yd;
#pragma unroll
    for (int p = S - 1; p >= 0; --p) {
      const double q = floor((y + 128.0) * 0x1.0p-8);
      const int t = (int)(y - 256.0 * q);  // in [-128, 127]
      pk |= (uint64_t)(uint8_t)(int8_t)t << (8 * p);
      y = q;
    }
  } else {
    long long y = (long long)yd;
#pragma unroll
    for (int p = S - 1; p >= 0; --p) {
      const long long t = ((y + 128) & 255) - 128;
      pk |= (uint64_t)(uint8_t)(int8_t)t << (8 * p);
      y = (y - t) >> 8;  // exact: y - t is a multiple of 256
    }
  }
  return pk;
}

// Byte p of each of 16 packed digit sets, in order, as one 16-byte slot.
__device__ __forceinline__ int4_t split_slot(const uint64_t *v, int p) {
  int4_t r;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    uint32_t word = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) word |= (uint32_t)((v[4 * w + b] >> (8 * p)) & 255u) << (8 * b);
    r[w] = (int)word;
  }
  return r;
}

// The real operand of part t of a complex value: re, im, re + im.
__device__ __forceinline__ double part_of(c128 z, int t) { return t == 0 ? z.re : t == 1 ? z.im : z.re + z.im; }

// Scales of the three parts of a row or column: exponents at expo[t * n + i],
// recombination factors at back[t * n + i].
struct ScaleOut {
  int *expo;
  double *back;
  uint64_t n;
};

// One block per complex row of X (A, or B^T when B is given N x K), all
// three parts.  The row is read once into LDS (rows of up to kRowTile
// elements, every thread's loads issued together; longer rows are read
// twice: their maxima first, then tile by tile) while the parts' largest
// magnitudes are found; then each thread turns 16 consecutive elements
// (16 bytes of depth) into S digits each and writes one 16-byte slot per
// part and plane.  LDS slot of element k: k ^ ((k / 16) % 16), so that the
// 16-element reads of consecutive threads and the coalesced writes both
// spread over the banks.
constexpr int kRowTile = 4096;
constexpr int kRowLoads = kRowTile / kThreads;

__device__ __forceinline__ uint32_t row_slot(uint32_t k) { return k ^ ((k >> 4) & 15u); }

template <int S>
__global__ void __launch_bounds__(kThreads) split_rows_kernel(const c128 *__restrict__ X, uint64_t rows, uint64_t K,
                                                              int8_t *__restrict__ out, uint64_t plane,
                                                              uint64_t part_bytes, uint64_t Kp, ScaleOut sc_out) {
  __shared__ c128 tile[kRowTile];
  __shared__ double red[kParts][kThreads / 64];
  const uint64_t row = blockIdx.x;
  if (row >= rows) return;  // uniform per block
  const c128 *x = X + row * K;
  const bool one_tile = K <= (uint64_t)kRowTile;
  double m[kParts] = {0.0, 0.0, 0.0};
  for (uint64_t k0 = 0; k0 < K; k0 += kRowTile) {
    c128 v[kRowLoads];
#pragma unroll
    for (int i = 0; i < kRowLoads; ++i) {
      const uint64_t k = k0 + threadIdx.x + (uint64_t)i * kThreads;
      v[i] = k < K ? x[k] : c128{0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < kRowLoads; ++i) {
      if (one_tile) tile[row_slot(threadIdx.x + i * kThreads)] = v[i];
#pragma unroll
      for (int t = 0; t < kParts; ++t) m[t] = fmax(m[t], fabs(part_of(v[i], t)));
    }
  }
#pragma unroll
  for (int t = 0; t < kParts; ++t) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m[t] = fmax(m[t], __shfl_xor(m[t], o));
    if ((threadIdx.x & 63) == 0) red[t][threadIdx.x >> 6] = m[t];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kParts; ++t) {
    m[t] = red[t][0];
#pragma unroll
    for (int w = 1; w < kThreads / 64; ++w) m[t] = fmax(m[t], red[t][w]);
  }
  const RowScale<S> sc[kParts] = {RowScale<S>(m[0]), RowScale<S>(m[1]), RowScale<S>(m[2])};
  if (threadIdx.x < kParts) {
    sc_out.expo[threadIdx.x * sc_out.n + row] = sc[threadIdx.x].e;
    sc_out.back[threadIdx.x * sc_out.n + row] = sc[threadIdx.x].back;
  }
  for (uint64_t t0 = 0; t0 < K; t0 += kRowTile) {
    const uint32_t tn = (uint32_t)std::min<uint64_t>(kRowTile, K - t0);
    if (!one_tile) {
      __syncthreads();  // the previous tile's readers are done
      c128 v[kRowLoads];
#pragma unroll
      for (int i = 0; i < kRowLoads; ++i) {
        const uint32_t k = threadIdx.x + i * kThreads;
        v[i] = k < tn ? x[t0 + k] : c128{0.0, 0.0};
      }
#pragma unroll
      for (int i = 0; i < kRowLoads; ++i) tile[row_slot(threadIdx.x + i * kThreads)] = v[i];
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < (tn + 15) / 16; c += kThreads) {
      c128 z[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) z[r] = 16 * c + r < tn ? tile[16 * c + (r ^ (c & 15u))] : c128{0.0, 0.0};
      const uint64_t depth = t0 + 16 * (uint64_t)c;
#pragma unroll
      for (int t = 0; t < kParts; ++t) {
        uint64_t v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = split_digits<S>(part_of(z[r], t), sc[t]);
#pragma unroll
        for (int p = 0; p < S; ++p)
          *reinterpret_cast<int4_t *>(out + t * part_bytes + (uint64_t)p * plane + tiled_at(row, depth, Kp)) =
              split_slot(v, p);
      }
    }
  }
}

// B read in place as K x N (row-major, the FP64 GEMM's layout), giving the
// planes split_rows_kernel gives for the rows of B^T, with no transposed copy
// of B.  First the columns' largest part magnitudes, by atomic maximum over
// blocks of rows on the bit patterns of non-negative doubles (their order is
// that of the integers), from zero.
constexpr int kColRows = 64;
constexpr int kColBatch = 16;  // loads in flight per thread

__global__ void zero_u64_kernel(unsigned long long *__restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

__global__ void __launch_bounds__(kThreads) col_max_kernel(const c128 *__restrict__ B, uint64_t K, uint64_t N,
                                                           unsigned long long *__restrict__ cmax) {
  // Block: kThreads consecutive columns (each row read as one contiguous
  // 4 KiB run by its four waves) x kColRows rows, kColBatch loads in flight
  // per thread; one atomic per part and column.
  const uint64_t j = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint64_t k0 = (uint64_t)blockIdx.y * kColRows;
  if (j >= N) return;  // no barriers here
  double m[kParts] = {0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < kColRows; i0 += kColBatch) {
    c128 v[kColBatch];
#pragma unroll
    for (int i = 0; i < kColBatch; ++i) {
      const uint64_t k = k0 + (uint64_t)(i0 + i);
      v[i] = k < K ? B[k * N + j] : c128{0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < kColBatch; ++i)
#pragma unroll
      for (int t = 0; t < kParts; ++t) m[t] = fmax(m[t], fabs(part_of(v[i], t)));
  }
#pragma unroll
  for (int t = 0; t < kParts; ++t)
    if (m[t] > 0.0) atomicMax(cmax + t * N + j, (unsigned long long)__double_as_longlong(m[t]));
}

// Digits of B (K x N) into the B^T planes of the three parts.  A block takes
// 64 columns (four row groups of the planes) x 64 rows of B (64 bytes of
// depth: one chunk column).  Thread (column j, rows 16 c .. 16 c + 15) reads
// its 16 elements (consecutive threads take consecutive columns: each load is
// a coalesced row segment) and puts the 16-byte slot of row j of every part
// and plane into an LDS copy of the block's chunks, laid out as in the
// planes; then every wave instruction stores one whole 1 KiB chunk.
template <int S>
__global__ void __launch_bounds__(kThreads) split_cols_kernel(const c128 *__restrict__ B, uint64_t K, uint64_t N,
                                                              const unsigned long long *__restrict__ cmax,
                                                              int8_t *__restrict__ out, uint64_t plane,
                                                              uint64_t part_bytes, uint64_t Kp, ScaleOut sc_out) {
  __shared__ int4_t chunk[kParts][S][64][4];
  const int jl = threadIdx.x & 63, c = threadIdx.x >> 6;
  const uint64_t j = (uint64_t)blockIdx.x * 64 + jl;
  const uint64_t k0 = (uint64_t)blockIdx.y * 64 + 16 * c;
  double m[kParts] = {0.0, 0.0, 0.0};
  if (j < N)
#pragma unroll
    for (int t = 0; t < kParts; ++t) m[t] = __longlong_as_double((long long)cmax[t * N + j]);
  const RowScale<S> sc[kParts] = {RowScale<S>(m[0]), RowScale<S>(m[1]), RowScale<S>(m[2])};
  if (j < N && blockIdx.y == 0 && c == 0)
#pragma unroll
    for (int t = 0; t < kParts; ++t) {
      sc_out.expo[t * sc_out.n + j] = sc[t].e;
      sc_out.back[t * sc_out.n + j] = sc[t].back;
    }
  c128 z[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) z[r] = (j < N && k0 + r < K) ? B[(k0 + r) * N + j] : c128{0.0, 0.0};
  // Physical slot of logical slot c in row j of a chunk: c ^ ((j / 2) % 4).
  const int phys = c ^ ((jl >> 1) & 3);
#pragma unroll
  for (int t = 0; t < kParts; ++t) {
    uint64_t v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = split_digits<S>(part_of(z[r], t), sc[t]);
#pragma unroll
    for (int p = 0; p < S; ++p) chunk[t][p][jl][phys] = split_slot(v, p);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t groups = Kp >> 6;  // 1 KiB chunks per row group of a plane
  constexpr int kChunks = kParts * S * 4;
#pragma unroll
  for (int ii = 0; ii < (kChunks + kThreads / 64 - 1) / (kThreads / 64); ++ii) {
    const int i = wave + ii * (kThreads / 64);
    if (i >= kChunks) break;
    const int t = i / (4 * S), p = (i / 4) % S, g = i & 3;
    const uint64_t row_group = (uint64_t)blockIdx.x * 4 + g;
    int8_t *dst = out + t * part_bytes + (uint64_t)p * plane + ((row_group * groups + blockIdx.y) << 10) + 16 * lane;
    *reinterpret_cast<int4_t *>(dst) = chunk[t][p][16 * g + (lane >> 2)][lane & 3];
  }
}

// Block tile: 8 waves in a WGM x WGN grid, each owning WM x WN outputs:
// 128 x 128 with 64 x 32 wave tiles for S <= 5, 128 x 64 with 32 x 32 wave
// tiles above.  Operand panels (S slices x rows x 64 bytes of depth) go from
// global memory straight into LDS (global_load_lds_dwordx4: each wave
// instruction fills 1 KiB = 16 rows x 64 bytes, the lanes choosing source
// addresses so that the 16-byte chunks land XOR-swizzled by row and the
// fragment reads hit distinct banks).  Two LDS stages when they fit: the next
// depth step streams in while the current one feeds the matrix cores.
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

// One launch of the GEMM kernels: parts part0 .. part0 + nparts - 1 of one
// depth chunk.  Part t reads planes As + t a_part, Bs + t b_part (this
// chunk's depth starts at the bases) and the scales of part t.  Products go
// to T (T1 at 0, T2 at M N, T3 at 2 M N, row-major M x N doubles), added
// to what is there when accumulate is set; with combine set, part 3 instead
// writes C = (T1 - T2, (T3 - T1) - T2).
struct OzArgs {
  const int8_t *As, *Bs;
  uint64_t a_plane, b_plane, a_part, b_part;
  uint64_t lda, ldb;  // plane depth in bytes (row length)
  const int *ea, *eb;
  const double *fa, *fb;  // recombination factors
  double *T;
  double2_t *C;
  uint64_t M, N, Kc;  // Kc: depth of this chunk (a multiple of BK)
  unsigned tilesM, tilesN;
  int part0, accumulate, combine;
};

// Tile and part of this workgroup.  Workgroups are dealt round-robin to the 8
// XCDs; give each XCD a contiguous range of (part, tile) indices, tiles walked
// in groups of kGroupM tile rows, so the blocks an XCD runs together share A
// and B panels in its L2.
struct OzTile {
  int part;
  uint64_t m0, n0;
};
__device__ __forceinline__ OzTile oz_tile(const OzArgs &a, int BM, int BN) {
  const unsigned T = gridDim.x, b = blockIdx.x;
  const unsigned xcd = b & 7, per = T >> 3, rem = T & 7;
  unsigned t = xcd * per + min(xcd, rem) + (b >> 3);
  const unsigned tiles = a.tilesM * a.tilesN;
  OzTile r;
  r.part = a.part0 + (int)(t / tiles);
  t %= tiles;
  const unsigned gsize = kGroupM * a.tilesN;
  const unsigned first_m = (t / gsize) * kGroupM;
  const unsigned gm_cnt = min((unsigned)kGroupM, a.tilesM - first_m);
  const unsigned in = t % gsize;
  r.m0 = (uint64_t)(first_m + in % gm_cnt) * BM;
  r.n0 = (uint64_t)(in / gm_cnt) * BN;
  return r;
}

// Recombines one output of part t from its S diagonal sums (row scale
// 2^er fr, column scale 2^ec fc) and stores it.
template <int S>
__device__ __forceinline__ void oz_store(const OzArgs &a, int t, uint64_t gm, uint64_t gn, const int *accd, double fr,
                                         int er, double fc, int ec) {
  double sum = 0.0;
#pragma unroll
  for (int d = S - 1; d >= 0; --d) sum += ldexp((double)accd[d], 8 * (2 * S - 2 - d));
  const double v = ldexp(sum * (fr * fc), er + ec);
  const uint64_t MN = a.M * a.N, e = gm * a.N + gn;
  if (t == 2 && a.combine) {
    double t3 = v;
    if (a.accumulate) t3 += a.T[2 * MN + e];  // T3 exists when the depth takes several chunks
    const double t1 = a.T[e], t2 = a.T[MN + e];
    a.C[e] = double2_t{t1 - t2, (t3 - t1) - t2};
  } else {
    double *dst = a.T + (uint64_t)t * MN + e;
    *dst = a.accumulate ? *dst + v : v;
  }
}

// 16x16 i32 result layout: lane l holds column l % 16 of rows 4 (l / 16) + v.
// The scales of the lane's 4 FA rows and FB columns are read once (the
// stores in between would otherwise make the compiler reload them per
// output: 4 scattered loads per output).
template <int S, int FA, int FB>
__device__ __forceinline__ void oz_epilogue(const OzArgs &a, const OzTile &tl, int wm, int wn, int lane,
                                            const int4_t (&acc)[S][FA][FB]) {
  const int t = tl.part;
  double fr[FA][4], fc[FB];
  int er[FA][4], ec[FB];
#pragma unroll
  for (int i = 0; i < FA; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const uint64_t gm = tl.m0 + wm + 16 * i + 4 * (lane >> 4) + v;
      const uint64_t r = (uint64_t)t * a.M + (gm < a.M ? gm : 0);
      fr[i][v] = a.fa[r];
      er[i][v] = a.ea[r];
    }
#pragma unroll
  for (int j = 0; j < FB; ++j) {
    const uint64_t gn = tl.n0 + wn + 16 * j + (lane & 15);
    const uint64_t c = (uint64_t)t * a.N + (gn < a.N ? gn : 0);
    fc[j] = a.fb[c];
    ec[j] = a.eb[c];
  }
#pragma unroll
  for (int i = 0; i < FA; ++i)
#pragma unroll
    for (int j = 0; j < FB; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint64_t gm = tl.m0 + wm + 16 * i + 4 * (lane >> 4) + v;
        const uint64_t gn = tl.n0 + wn + 16 * j + (lane & 15);
        if (gm < a.M && gn < a.N) {
          int accd[S];
#pragma unroll
          for (int d = 0; d < S; ++d) accd[d] = acc[d][i][j][v];
          oz_store<S>(a, t, gm, gn, accd, fr[i][v], er[i][v], fc[j], ec[j]);
        }
      }
}

template <int S, int WM, int WN, int WGM, int WGN, int MINW>
__global__ void __launch_bounds__(64 * WGM * WGN, MINW) ozaki_gemm_kernel(OzArgs a) {
  using Z = OzShape<S, WM, WN, WGM, WGN>;
  constexpr int BM = Z::BM, BN = Z::BN, NST = Z::NST;
  __shared__ int4_t sA[NST][S * BM * 4];
  __shared__ int4_t sB[NST][S * BN * 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;
  const OzTile tl = oz_tile(a, BM, BN);
  const int8_t *As = a.As + tl.part * a.a_part, *Bs = a.Bs + tl.part * a.b_part;
  const uint64_t m0 = tl.m0, n0 = tl.n0;

  int4_t acc[S][Z::FA][Z::FB];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < Z::FA; ++i)
#pragma unroll
      for (int j = 0; j < Z::FB; ++j) acc[d][i][j] = int4_t{0, 0, 0, 0};

  // A chunk of the tiled planes is one wave instruction: lane i copies its
  // 16-byte slot i into LDS slot i.  The planes hold whole tiles of rows (the
  // rows past M or N are padding whose products land only in outputs that are
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
        const int8_t *base = isA ? As + p * a.a_plane + ((m0 >> 4) + g) * (a.lda << 4) + (kt << 10)
                                 : Bs + p * a.b_plane + ((n0 >> 4) + g) * (a.ldb << 4) + (kt << 10);
        int4_t *dst = isA ? &sA[st][(p * rows + g * 16) * 4] : &sB[st][(p * rows + g * 16) * 4];
        __builtin_amdgcn_global_load_lds((const void *)(base + off), (void __attribute__((address_space(3))) *)dst, 16,
                                         0, 0);
      }
    }
  };

  // The caller pads the depth to a multiple of BK with zeros.
  const uint64_t nk = a.Kc / BK;
  issue(0, 0);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (uint64_t kt = 0; kt < nk; ++kt) {
    const int cur = (NST == 2) ? (int)(kt & 1) : 0;
    // Fragments (16x16x64 i8): lane l holds row / column l % 16, k bytes 16 (l / 16) .. +15.
    // All B slices stay in registers; A slices stream through, one at a time.
    int4_t fb[S][Z::FB];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int j = 0; j < Z::FB; ++j) fb[q][j] = sB[cur][lds_at(q, wn + 16 * j + (lane & 15), lane >> 4, BN)];
    // The next step streams in during the MFMAs below (issued after this
    // step's B fragment reads: measured 1% faster than before them).
    if (NST == 2 && kt + 1 < nk) issue(cur ^ 1, kt + 1);
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
  oz_epilogue<S, Z::FA, Z::FB>(a, tl, wm, wn, lane, acc);
}

// Register-staged variant for S >= 6 (128 x 64 tiles, 8 waves of 32 x 32):
// each thread loads its 16-byte chunks of the next depth step into registers
// before the MFMAs of the current one and writes them to the other LDS stage
// after them.  Same arithmetic and result layout as ozaki_gemm_kernel.
template <int S>
__global__ void __launch_bounds__(512) ozaki_gemm_staged(OzArgs a) {
  constexpr int BM = 128, BN = 64, kGemmThreads = 512;
  constexpr int NST = (S <= 6) ? 2 : 1;
  __shared__ int4_t sA[NST][S * BM * 4];
  __shared__ int4_t sB[NST][S * BN * 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const OzTile tl = oz_tile(a, BM, BN);
  const int8_t *As = a.As + tl.part * a.a_part, *Bs = a.Bs + tl.part * a.b_part;
  const uint64_t m0 = tl.m0, n0 = tl.n0;

  int4_t acc[S][2][2];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[d][i][j] = int4_t{0, 0, 0, 0};

  // Staging: A chunk (row tid / 4, 16-byte chunk tid % 4) of every slice; B
  // chunks tid + 512 i over the S x 64 x 4 chunks of the B panel.  Rows past
  // M or N are padding rows of the planes (read, never stored).
  constexpr int BL = (S * BN * 4 + kGemmThreads - 1) / kGemmThreads;
  const int4_t zero = {0, 0, 0, 0};
  int4_t pa[S], pb[BL];
  const int ar = tid >> 2, ac = tid & 3;
  auto fetch = [&](uint64_t k0) {
    const uint64_t ga = m0 + ar;
#pragma unroll
    for (int p = 0; p < S; ++p)
      pa[p] = *reinterpret_cast<const int4_t *>(As + p * a.a_plane + tiled_at(ga, k0 + ac * 16, a.lda));
#pragma unroll
    for (int i = 0; i < BL; ++i) {
      const int c = tid + kGemmThreads * i;
      const int p = c / (BN * 4), r = (c % (BN * 4)) >> 2, ch = c & 3;
      pb[i] = c < S * BN * 4
                  ? *reinterpret_cast<const int4_t *>(Bs + p * a.b_plane + tiled_at(n0 + r, k0 + ch * 16, a.ldb))
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

  const uint64_t nk = a.Kc / BK;
  fetch(0);
  stash(0);
  __syncthreads();
  for (uint64_t kt = 0; kt < nk; ++kt) {
    const int cur = (NST == 2) ? (int)(kt & 1) : 0;
    if (kt + 1 < nk) fetch((kt + 1) * BK);  // in flight during the MFMAs below
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
  oz_epilogue<S, 2, 2>(a, tl, wm, wn, lane, acc);
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
void launch_ozaki_shape(hipStream_t st, OzArgs a, int nparts) {
  using Z = OzShape<S, WM, WN, WGM, WGN>;
  a.tilesM = (unsigned)((a.M + Z::BM - 1) / Z::BM);
  a.tilesN = (unsigned)((a.N + Z::BN - 1) / Z::BN);
  hipLaunchKernelGGL((ozaki_gemm_kernel<S, WM, WN, WGM, WGN, MINW>), dim3(a.tilesM * a.tilesN * nparts),
                     dim3(Z::THREADS), 0, st, a);
}

template <int S>
void launch_ozaki(hipStream_t st, const OzArgs &a, int nparts) {
  if constexpr (S >= 6) {
    OzArgs b = a;
    b.tilesM = (unsigned)((a.M + 127) / 128);
    b.tilesN = (unsigned)((a.N + 63) / 64);
    hipLaunchKernelGGL((ozaki_gemm_staged<S>), dim3(b.tilesM * b.tilesN * nparts), dim3(512), 0, st, b);
  } else {
    switch (g_split_tile) {
      case 1: launch_ozaki_shape<S, 32, 32, 2, 2, 2>(st, a, nparts); break;
      case 2: launch_ozaki_shape<S, 64, 32, 2, 2, 1>(st, a, nparts); break;
      case 3: launch_ozaki_shape<S, 32, 16, 2, 4, 4>(st, a, nparts); break;
      default: launch_ozaki_shape<S, 64, 32, 2, 4, 1>(st, a, nparts); break;
    }
  }
}

// A's planes on `stream`, B's on `side`: independent memory-bound passes
// that run together (forked from and joined back into `stream` by the two
// events, which a graph capture of `stream` follows).
template <int S>
hipError_t split_planes(hipStream_t stream, hipStream_t side, hipEvent_t fork, hipEvent_t join, const c128 *A,
                        const c128 *B, bool b_cols, uint64_t M, uint64_t N, uint64_t K, int8_t *As, uint64_t a_plane,
                        int8_t *Bs, uint64_t b_plane, uint64_t Kp, ScaleOut sa, ScaleOut sb,
                        unsigned long long *cmax) {
  hipError_t e = hipEventRecord(fork, stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(side, fork, 0);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((split_rows_kernel<S>), dim3((unsigned)M), dim3(kThreads), 0, stream, A, M, K, As, a_plane,
                     (uint64_t)S * a_plane, Kp, sa);
  if (b_cols) {
    // B as K x N: the columns' maxima, then digits read in place.
    hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)std::min<uint64_t>((kParts * N + 255) / 256, 65536)), dim3(256),
                       0, side, cmax, kParts * N);
    hipLaunchKernelGGL(col_max_kernel, dim3((unsigned)((N + kThreads - 1) / kThreads),
                                            (unsigned)((K + kColRows - 1) / kColRows)),
                       dim3(kThreads), 0, side, B, K, N, cmax);
    hipLaunchKernelGGL((split_cols_kernel<S>), dim3((unsigned)((N + 63) / 64), (unsigned)((K + 63) / 64)),
                       dim3(kThreads), 0, side, B, K, N, cmax, Bs, b_plane, (uint64_t)S * b_plane, Kp, sb);
  } else {
    // B^T as N x K: one block per row.
    hipLaunchKernelGGL((split_rows_kernel<S>), dim3((unsigned)N), dim3(kThreads), 0, side, B, N, K, Bs, b_plane,
                       (uint64_t)S * b_plane, Kp, sb);
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipEventRecord(join, side);
  if (e == hipSuccess) e = hipStreamWaitEvent(stream, join, 0);
  return e;
}

// The depth chunks of one split GEMM: parts 1 and 2 in one launch, then part
// 3, which on the last chunk also forms C.
template <int S>
hipError_t split_gemm(qtnh_ctx_t ctx, const c128 *A, const c128 *B, bool b_cols, c128 *C, uint64_t M, uint64_t N,
                      uint64_t K, int8_t *As, uint64_t a_plane, int8_t *Bs, uint64_t b_plane, uint64_t Kp, ScaleOut sa,
                      ScaleOut sb, unsigned long long *cmax, double *T, uint64_t chunk) {
  hipStream_t stream = ctx->stream;
  const hipError_t e = split_planes<S>(stream, ctx->side, ctx->ev_fork, ctx->ev_join, A, B, b_cols, M, N, K, As,
                                       a_plane, Bs, b_plane, Kp, sa, sb, cmax);
  if (e != hipSuccess) return e;
  OzArgs a;
  memset(&a, 0, sizeof(a));
  a.a_plane = a_plane;
  a.b_plane = b_plane;
  a.a_part = (uint64_t)S * a_plane;
  a.b_part = (uint64_t)S * b_plane;
  a.lda = Kp;
  a.ldb = Kp;
  a.ea = sa.expo;
  a.eb = sb.expo;
  a.fa = sa.back;
  a.fb = sb.back;
  a.T = T;
  a.C = reinterpret_cast<double2_t *>(C);
  a.M = M;
  a.N = N;
  // Depth chunk k0 (a multiple of 64) starts at chunk k0 / 64 of every row
  // group of the tiled planes: byte 16 k0.
  for (uint64_t k0 = 0; k0 < Kp; k0 += chunk) {
    a.Kc = std::min(chunk, Kp - k0);
    a.As = As + 16 * k0;
    a.Bs = Bs + 16 * k0;
    a.accumulate = k0 > 0;
    a.part0 = 0;
    a.combine = 0;
    launch_ozaki<S>(stream, a, 2);
    a.part0 = 2;
    a.combine = k0 + chunk >= Kp;
    launch_ozaki<S>(stream, a, 1);
  }
  return hipGetLastError();
}

// Rows of a slice plane: whole 128-row tiles (every block tile is at most 128
// rows of A and 128 rows of B^T), so the GEMM kernels read padding rows past
// the end instead of clamping row indices.  The padding rows are never
// written; their products reach only outputs past M or N, which are not stored.
uint64_t plane_rows(uint64_t rows) { return (rows + 127) / 128 * 128; }

// int32 accumulators stay exact while (S pairs) * 128 * 128 * depth < 2^31:
// depth chunks of 2^13 (S <= 8).
constexpr uint64_t kSplitChunk = 8192;

// Workspace of a split GEMM: the planes of the three parts of A and B, T1
// and T2 (and T3 when the depth takes several chunks), the scales and the
// columns' maxima, each region 256-byte aligned.
struct SplitLayout {
  uint64_t Kp, a_plane, b_plane, offB, offT, offS, bytes;
};
uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }
SplitLayout split_layout(int S, uint64_t M, uint64_t N, uint64_t K) {
  SplitLayout L;
  L.Kp = (K + BK - 1) / BK * BK;
  L.a_plane = plane_rows(M) * L.Kp;
  L.b_plane = plane_rows(N) * L.Kp;
  L.offB = align256((uint64_t)kParts * S * L.a_plane);
  L.offT = L.offB + align256((uint64_t)kParts * S * L.b_plane);
  const uint64_t nt = L.Kp > kSplitChunk ? 3 : 2;
  L.offS = L.offT + align256(nt * M * N * sizeof(double));
  L.bytes = L.offS + align256(kParts * (M + N) * sizeof(int)) + align256(kParts * (M + N) * sizeof(double)) +
            align256(kParts * N * sizeof(unsigned long long));
  return L;
}

}  // namespace

// Number of 8-bit slices of the split-precision GEMM (default 5).
static int g_split_slices = 5;

extern "C" void qtnh_tune_split(int slices) { g_split_slices = std::max(1, std::min(slices, kMaxSlices)); }

extern "C" void qtnh_tune_split_tile(int shape) { g_split_tile = (shape >= 0 && shape <= 3) ? shape : 0; }

int split_slices() { return g_split_slices; }

// A-priori componentwise bound.  Per real product T = P Q of depth K, with
// m_i, n_j the largest magnitudes of P's row i and Q's column j and
// c = 127 * 256^(S-1): rounding Y leaves |rho| <= 1/2 per element
// (K m n (1/c + 1/(4c^2)) in all), and the dropped diagonals p + q >= S, digits
// of at most 128, add at most K m n (2^14 256^(S-2) S 256/255) / c^2, together
// at most K m n (S + 3) 256^-S.  With a_i = max_k max(|Re A_ik|, |Im A_ik|),
// b_j likewise for B: m, n <= a, b for T1, T2 and <= 2a, 2b for T3, so
// |dRe C| <= 2, |dIm C| <= 6 times K (S + 3) 256^-S a_i b_j, and
//   |C_ij - C~_ij| <= eta a_i b_j,  eta = 6.4 K (S + 3) 256^-S
// (6.4 > sqrt(2^2 + 6^2) also covers the binary64 rounding of Ar + Ai and of
// the recombination, below K 2^-50 a b for S <= 6; a result that underflows
// is rounded twice and may differ by subnormal spacings, 2^-1074 each).
double split_error_bound(int S, uint64_t K) { return 6.4 * (double)(S + 3) * ldexp(1.0, -8 * S) * (double)K; }

qtnh_status launch_zgemm_split(qtnh_ctx_t ctx, const c128 *A, const c128 *B, bool b_cols, c128 *C, uint64_t M,
                               uint64_t N, uint64_t K, void *scratch, uint64_t scratch_bytes, double *err_out) {
  const int S = g_split_slices;
  const SplitLayout L = split_layout(S, M, N, K);
  if (scratch_bytes < L.bytes) return set_error(ctx, QTNH_EINVAL, "split: scratch too small");
  if (M > 0x7fffffffull || N > 0x7fffffffull) return set_error(ctx, QTNH_EUNSUPPORTED, "split: too many rows");
  const uint64_t grid = ((M + 127) / 128) * ((N + 63) / 64) * 2;
  if (grid > 0x7fffffffull) return set_error(ctx, QTNH_EUNSUPPORTED, "split: grid too large");
  char *base = (char *)scratch;
  int8_t *As = (int8_t *)base;
  int8_t *Bs = (int8_t *)(base + L.offB);
  double *T = (double *)(base + L.offT);
  char *s = base + L.offS;
  int *ea = (int *)s;
  int *eb = ea + kParts * M;
  s += align256(kParts * (M + N) * sizeof(int));
  double *fa = (double *)s;
  double *fb = fa + kParts * M;
  s += align256(kParts * (M + N) * sizeof(double));
  unsigned long long *cmax = (unsigned long long *)s;
  const ScaleOut sa{ea, fa, M}, sb{eb, fb, N};
  // Padding columns of the slice planes must read as zero.
  if (L.Kp != K) {
    qtnh_status st = launch_zero(ctx, As, L.offT);
    if (st) return st;
  }
  hipError_t e = hipSuccess;
  switch (S) {
#define QTNH_SPLIT_CASE(s)                                                                                      \
  case s:                                                                                                      \
    e = split_gemm<s>(ctx, A, B, b_cols, C, M, N, K, As, L.a_plane, Bs, L.b_plane, L.Kp, sa, sb, cmax, T,     \
                      kSplitChunk);                                                                            \
    break;
    QTNH_SPLIT_CASE(1)
    QTNH_SPLIT_CASE(2)
    QTNH_SPLIT_CASE(3)
    QTNH_SPLIT_CASE(4)
    QTNH_SPLIT_CASE(5)
    QTNH_SPLIT_CASE(6)
    QTNH_SPLIT_CASE(7)
    QTNH_SPLIT_CASE(8)
#undef QTNH_SPLIT_CASE
    default:
      return set_error(ctx, QTNH_EINVAL, "split: %d slices", S);
  }
  qtnh_status st = check_hip(ctx, e, "split gemm launch");
  if (st) return st;
  if (err_out) *err_out = split_error_bound(S, K);
  return QTNH_OK;
}

uint64_t split_scratch_bytes(uint64_t M, uint64_t N, uint64_t K) { return split_layout(g_split_slices, M, N, K).bytes; }

}  // namespace qtnh_hip
