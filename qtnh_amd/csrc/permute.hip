// Axis permutation and axis swap for complex128 tensors.
//
// Semantics (reference src/tensor/dense.cpp:152-169, 354-488): the result's
// axis ptup[i] is the source's axis i.  Pure data movement, bit-exact.
//
// Power-of-two dims: the permutation is a permutation of index bits.  A
// workgroup moves one tile = the 2^a lowest source elements (contiguous in the
// source) x the source bits that land in the 2^b lowest result elements
// (contiguous in the result), staged through LDS, so that both the global
// reads and the global writes are runs of >= 2^min(a,b) * 16 bytes.  The LDS
// slot of tile element t is t ^ ((t >> a) & (2^a - 1)): reads in source order
// and writes in result order both spread over the banks.
//
// Other dims: one thread per result element with mixed-radix decoding.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

constexpr int kPermThreads = 256;

struct PermBitsArgs {
  uint64_t in_hi[64];   // source offset of the tile bits above a, indexed by t >> a
  uint64_t out_hi[64];  // result offset of the tile bits above b, indexed by s >> b
  uint16_t s2t_lo[64];  // tile index t of result-order index s, from s bits [0, 6)
  uint16_t s2t_hi[64];  //   ... and from s bits [6, 12)
  uint8_t ob_in[64];    // source bit of outer bit j
  uint8_t ob_out[64];   // result bit of outer bit j
  int a, b, u;          // source-low bits, result-low bits, tile bits
  int n_outer;
};

__device__ inline uint32_t lds_slot(uint32_t t, int a) { return t ^ ((t >> a) & ((1u << a) - 1u)); }

typedef double double2_t __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ inline void store_d2(c128 *p, double2_t v) {
  if (NT)
    __builtin_nontemporal_store(v, reinterpret_cast<double2_t *>(p));
  else
    *reinterpret_cast<double2_t *>(p) = v;
}

template <bool NT>
__device__ inline double2_t load_d2(const double2_t *p) {
  if (NT) return __builtin_nontemporal_load(p);
  return *p;
}

template <bool NT>
__device__ inline void store_c128(c128 *p, c128 v) {
  if (NT) {
    double2_t w = {v.re, v.im};
    __builtin_nontemporal_store(w, reinterpret_cast<double2_t *>(p));
  } else {
    *p = v;
  }
}

// Tile and outer offsets of tile g.
__device__ inline void tile_bases(const PermBitsArgs &args, uint64_t g, uint64_t &xb, uint64_t &yb) {
  xb = 0;
  yb = 0;
  for (int j = 0; j < args.n_outer; ++j) {
    const uint64_t bit = (g >> j) & 1ull;
    xb |= bit << args.ob_in[j];
    yb |= bit << args.ob_out[j];
  }
}

// Persistent blocks walk tiles of 2^U elements, EPT elements per thread with
// no per-element branches.  The loads of the next tile are issued before the
// current tile is stored, so reads and writes overlap inside each block.  A
// 2^12-element tile takes 64 KiB of LDS: at most two blocks (two waves per
// SIMD) are resident, which the register budget below assumes.
template <int U, typename Off, bool NT>
__global__ void __launch_bounds__(kPermThreads) __attribute__((amdgpu_waves_per_eu(1, 2)))
permute_bits_kernel(const c128 *__restrict__ src, c128 *__restrict__ dst, PermBitsArgs args, uint64_t n_tiles) {
  constexpr uint32_t T = 1u << U;
  constexpr int EPT = T >= kPermThreads ? (int)(T / kPermThreads) : 1;
  extern __shared__ c128 tile[];
  // Small tiles (T < threads): spare threads skip the data movement but keep
  // taking part in the barriers.
  const bool live = T >= kPermThreads || threadIdx.x < T;
  const uint32_t amask = (1u << args.a) - 1u;
  const uint32_t bmask = (1u << args.b) - 1u;
  // Per-thread tile offsets are the same for every tile: compute them once.
  // Narrow tensors (< 2^28 elements) keep 32-bit byte offsets against a
  // uniform base pointer; LDS slots of the load and store phases share a register.
  Off xo[EPT], yo[EPT];
  uint32_t slots[EPT];
  constexpr Off kScale = sizeof(Off) == 4 ? (Off)sizeof(c128) : (Off)1;
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    const uint32_t e = live ? threadIdx.x + r * kPermThreads : 0;
    xo[r] = (Off)((e & amask) | args.in_hi[e >> args.a]) * kScale;
    const uint32_t t = args.s2t_lo[e & 63u] | args.s2t_hi[e >> 6];
    yo[r] = (Off)((e & bmask) | args.out_hi[e >> args.b]) * kScale;
    slots[r] = lds_slot(e, args.a) | (lds_slot(t, args.a) << 16);
  }
  auto at = [](const c128 *base, Off off) -> const double2_t * {
    if (sizeof(Off) == 4) return reinterpret_cast<const double2_t *>(reinterpret_cast<const char *>(base) + off);
    return reinterpret_cast<const double2_t *>(base + off);
  };
  auto at_mut = [](c128 *base, Off off) -> c128 * {
    if (sizeof(Off) == 4) return reinterpret_cast<c128 *>(reinterpret_cast<char *>(base) + off);
    return base + off;
  };
  double2_t *ltile = reinterpret_cast<double2_t *>(tile);
  uint64_t g = blockIdx.x;
  if (g >= n_tiles) return;  // uniform: the whole block leaves
  uint64_t xb, yb;
  tile_bases(args, g, xb, yb);
  double2_t v[EPT];
  if (live) {
    const c128 *s = src + xb;
#pragma unroll
    for (int r = 0; r < EPT; ++r) v[r] = load_d2<NT>(at(s, xo[r]));
  }
  while (true) {
    if (live) {
#pragma unroll
      for (int r = 0; r < EPT; ++r) ltile[slots[r] & 0xffffu] = v[r];
    }
    __syncthreads();
    const uint64_t gn = g + gridDim.x;
    const bool more = gn < n_tiles;  // uniform
    uint64_t xn = 0, yn = 0;
    if (more) {
      tile_bases(args, gn, xn, yn);
      if (live) {
        const c128 *s = src + xn;
#pragma unroll
        for (int r = 0; r < EPT; ++r) v[r] = load_d2<NT>(at(s, xo[r]));
      }
    }
    if (live) {
      c128 *d = dst + yb;
#pragma unroll
      for (int r = 0; r < EPT; ++r) store_d2<NT>(at_mut(d, yo[r]), ltile[slots[r] >> 16]);
    }
    __syncthreads();
    if (!more) break;
    g = gn;
    yb = yn;
  }
}

// A square dense gate (2^NB terms, 2^NB outputs) contracted into a state with
// its outputs on the lowest result bits, fused into the bit permutation that
// carries the state's other bits to their result positions.  The permutation
// sends the gate's term bits to result bits 0 .. NB-1, so a tile holds whole
// result rows.  The load phase is permute_bits_kernel's; in the store phase a
// thread takes the same result elements as there (consecutive lanes,
// consecutive elements, so each store instruction writes whole lines), reads
// the 2^NB terms of the element's row from the tile (lanes of one row read the
// same slots) and forms the element: 0 + the sum over the terms in the
// reference order (term t = 0, 1, ...) of separately rounded complex products,
// contract_direct_rows's arithmetic bit for bit.  A thread's elements all have
// output index tid mod 2^NB, so it loads its 2^NB coefficients once.
template <int NB>
struct GateRowArgs {
  const c128 *B;
  uint32_t xb[(1 << NB) * (1 << NB)];  // B offset of term t for output o, at t 2^NB + o
};

template <int U, typename Off, int NB>
__global__ void __launch_bounds__(kPermThreads) __attribute__((amdgpu_waves_per_eu(1, 2)))
permute_gate_kernel(const c128 *__restrict__ src, c128 *__restrict__ dst, PermBitsArgs args, GateRowArgs<NB> gr,
                    uint64_t n_tiles) {
  constexpr uint32_t T = 1u << U;
  constexpr int N = 1 << NB;
  constexpr int EPT = (int)(T / kPermThreads);
  static_assert(T >= kPermThreads, "every thread takes elements of a tile");
  extern __shared__ c128 tile[];
  const uint32_t amask = (1u << args.a) - 1u;
  const uint32_t bmask = (1u << args.b) - 1u;
  constexpr Off kScale = sizeof(Off) == 4 ? (Off)sizeof(c128) : (Off)1;
  // Tile element e = tid + 256 r: its source offset, its LDS slot in the load
  // phase, its result offset, and the slot of its row's term 0 (the row's
  // result-order index e - o; term t sits at tile index s2t(e - o) XOR s2t(t),
  // and lds_slot is linear over XOR).
  Off xo[EPT], yo[EPT];
  uint32_t slots[EPT];
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    const uint32_t e = threadIdx.x + r * kPermThreads;
    xo[r] = (Off)((e & amask) | args.in_hi[e >> args.a]) * kScale;
    yo[r] = (Off)((e & bmask) | args.out_hi[e >> args.b]) * kScale;
    const uint32_t s = e & ~(uint32_t)(N - 1);
    slots[r] = lds_slot(e, args.a) | (lds_slot(args.s2t_lo[s & 63u] | args.s2t_hi[s >> 6], args.a) << 16);
  }
  uint32_t sterm[N];
  double2_t cf[N];
  const int o = (int)(threadIdx.x & (N - 1));
#pragma unroll
  for (int t = 0; t < N; ++t) {
    sterm[t] = lds_slot(args.s2t_lo[t], args.a);
    const c128 c = gr.B[gr.xb[t * N + o]];
    cf[t] = double2_t{c.re, c.im};
  }
  auto at = [](const c128 *base, Off off) -> const double2_t * {
    if (sizeof(Off) == 4) return reinterpret_cast<const double2_t *>(reinterpret_cast<const char *>(base) + off);
    return reinterpret_cast<const double2_t *>(base + off);
  };
  auto at_mut = [](c128 *base, Off off) -> c128 * {
    if (sizeof(Off) == 4) return reinterpret_cast<c128 *>(reinterpret_cast<char *>(base) + off);
    return base + off;
  };
  double2_t *ltile = reinterpret_cast<double2_t *>(tile);
  uint64_t g = blockIdx.x;
  if (g >= n_tiles) return;  // uniform: the whole block leaves
  uint64_t xb, yb;
  tile_bases(args, g, xb, yb);
  double2_t v[EPT];
  {
    const c128 *s = src + xb;
#pragma unroll
    for (int r = 0; r < EPT; ++r) v[r] = load_d2<true>(at(s, xo[r]));
  }
  while (true) {
#pragma unroll
    for (int r = 0; r < EPT; ++r) ltile[slots[r] & 0xffffu] = v[r];
    __syncthreads();
    const uint64_t gn = g + gridDim.x;
    const bool more = gn < n_tiles;  // uniform
    uint64_t xn = 0, yn = 0;
    if (more) {
      tile_bases(args, gn, xn, yn);
      const c128 *s = src + xn;
#pragma unroll
      for (int r = 0; r < EPT; ++r) v[r] = load_d2<true>(at(s, xo[r]));
    }
    c128 *d = dst + yb;
#pragma unroll
    for (int r = 0; r < EPT; ++r) {
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int t = 0; t < N; ++t) {
        const double2_t x = ltile[(slots[r] >> 16) ^ sterm[t]];
        double pr, pi;
        cmul_parts(x.x, x.y, cf[t].x, cf[t].y, &pr, &pi);
        ar = ar + pr;
        ai = ai + pi;
      }
      store_d2<true>(at_mut(d, yo[r]), double2_t{ar, ai});
    }
    __syncthreads();
    if (!more) break;
    g = gn;
    yb = yn;
  }
}

// General dims: result element y -> source offset.  nd: result dims, jump:
// source stride of the source axis placed at each result axis.
struct PermGenArgs {
  uint64_t nd[kMaxRank];
  uint64_t jump[kMaxRank];
  int rank;
};

__global__ void permute_general_kernel(const c128 *__restrict__ src, c128 *__restrict__ dst, PermGenArgs args,
                                       uint64_t n) {
  for (uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; y < n; y += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t rest = y, x = 0;
    for (int q = args.rank - 1; q >= 0; --q) {
      uint64_t d = args.nd[q];
      uint64_t i = rest % d;
      rest /= d;
      x += i * args.jump[q];
    }
    dst[y] = src[x];
  }
}

__global__ void copy_kernel(const c128 *__restrict__ src, c128 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Swap of two single-bit axes at bit positions lo < hi: exchange the elements
// with (bit hi, bit lo) = (1, 0) and (0, 1).  One thread per such pair.
__global__ void swap_bits_kernel(c128 *__restrict__ buf, int lo, int hi, uint64_t n_pairs) {
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_pairs;
       g += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t b = g;
    b = ((b >> lo) << (lo + 1)) | (b & ((1ull << lo) - 1));
    b = ((b >> hi) << (hi + 1)) | (b & ((1ull << hi) - 1));
    uint64_t e1 = b | (1ull << hi), e2 = b | (1ull << lo);
    c128 v1 = buf[e1], v2 = buf[e2];
    buf[e1] = v2;
    buf[e2] = v1;
  }
}

// Swap of axes with strides si, sj and common dim d: the thread of the
// element whose value at axis i is below its value at axis j exchanges it
// with the transposed element.
__global__ void swap_general_kernel(c128 *__restrict__ buf, uint64_t n, uint64_t si, uint64_t sj, uint64_t d) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t vi = (e / si) % d, vj = (e / sj) % d;
    if (vi < vj) {
      uint64_t o = e + (vj - vi) * si - (vj - vi) * sj;
      c128 v1 = buf[e], v2 = buf[o];
      buf[e] = v2;
      buf[o] = v1;
    }
  }
}

}  // namespace

// Tunables of the bit-permute tile.  a / b: starting source-low / result-low
// bits; the tile then grows until it holds u_target bits (<= 12, 64 KiB of
// LDS).  order bit 0: walk tiles by source bit position (0) or by the larger
// of the source and result positions (1, page locality on both sides).
// order bit 1: fill the tile up to u_target with outer bits whose source and
// result positions lie near 2^10 elements (runs of 2^a and 2^b elements at
// strides of 16 KiB or so), instead of growing a and b (longer runs).
// order bit 2: walk tiles by the distance of the outer bits' source and
// result positions from 10 (tiles in flight together differ there).  nt:
// loads and stores bypass the caches.  blocks_per_cu: blocks per compute unit,
// each looping over tiles.
struct PermTune {
  int a = 6, b = 6, u_target = 12, order = 1, nt = 1, blocks_per_cu = 2;
};
static PermTune g_perm;

// Per-permutation tuning.  No one tile shape is best for every permutation on
// MI355X (rank-28, TB/s: the default 12-bit tile with two blocks per CU gives
// 4.5-5.3; an 11-bit tile with 64-128 blocks per CU, each looping over a few
// tiles, gives 5.9-6.0 on some permutations and 4.7 on others; where the
// natural tile is short, filling it with outer bits near 2^10 gives 6.0-6.2
// against 5.4-5.7 for longer runs), and every shape writes the same bits.  So the first permute of a
// large state (>= 2^23 elements) with a new (bit count, bit permutation)
// signature on a device times each candidate below on the caller's buffers
// and keeps the fastest for that signature.  The timing synchronises the
// stream once; it is skipped (defaults used, nothing cached) while the stream
// is being captured into a graph.
static bool g_perm_auto = true;
static const PermTune kPermCandidates[] = {{6, 6, 12, 1, 1, 2},  {6, 6, 12, 1, 1, 128}, {5, 6, 11, 0, 1, 64},
                                           {5, 6, 11, 1, 1, 128}, {5, 6, 11, 1, 1, 32},  {6, 5, 11, 0, 1, 256},
                                           {6, 6, 12, 2, 1, 128}, {6, 6, 11, 3, 1, 64},  {6, 6, 12, 3, 1, 2},
                                           {5, 5, 11, 2, 1, 128}, {6, 5, 11, 2, 1, 128}, {6, 6, 12, 4, 1, 128},
                                           {5, 5, 11, 6, 1, 128}};
constexpr int kPermTimedRuns = 3;
constexpr uint64_t kPermAutoMin = 1ull << 23;
constexpr size_t kPermCacheMax = 4096;
static std::mutex g_perm_mu;
static std::map<std::string, PermTune> g_perm_cache;

extern "C" void qtnh_tune_permute(int a, int b, int u_target, int order, int nt, int blocks_per_cu) {
  g_perm.a = std::max(1, std::min(a, 12));
  g_perm.b = std::max(1, std::min(b, 12));
  g_perm.u_target = std::max(1, std::min(u_target, 12));
  g_perm.order = order;
  g_perm.nt = nt;
  g_perm.blocks_per_cu = std::max(1, blocks_per_cu);
  g_perm_auto = false;
}

extern "C" void qtnh_tune_permute_auto(int on) {
  std::lock_guard<std::mutex> lock(g_perm_mu);
  g_perm_auto = on != 0;
  if (on) g_perm = PermTune{};
  g_perm_cache.clear();
}

static qtnh_status launch_copy(qtnh_ctx_t ctx, const c128 *src, c128 *dst, uint64_t n) {
  unsigned blocks = grid_for(n, 256, 65536);
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, ctx->stream, src, dst, n);
  return check_hip(ctx, hipGetLastError(), "copy launch");
}

namespace {
struct PermLaunch {
  hipStream_t stream;
  unsigned blocks;
  size_t lds;
  const c128 *src;
  c128 *dst;
  uint64_t n_tiles;
  bool narrow, nt;
};

template <int U>
void launch_tile(const PermLaunch &L, const PermBitsArgs &args) {
  if (L.narrow && !L.nt)
    hipLaunchKernelGGL((permute_bits_kernel<U, uint32_t, false>), dim3(L.blocks), dim3(kPermThreads), L.lds, L.stream,
                       L.src, L.dst, args, L.n_tiles);
  else if (L.narrow)
    hipLaunchKernelGGL((permute_bits_kernel<U, uint32_t, true>), dim3(L.blocks), dim3(kPermThreads), L.lds, L.stream,
                       L.src, L.dst, args, L.n_tiles);
  else if (!L.nt)
    hipLaunchKernelGGL((permute_bits_kernel<U, uint64_t, false>), dim3(L.blocks), dim3(kPermThreads), L.lds, L.stream,
                       L.src, L.dst, args, L.n_tiles);
  else
    hipLaunchKernelGGL((permute_bits_kernel<U, uint64_t, true>), dim3(L.blocks), dim3(kPermThreads), L.lds, L.stream,
                       L.src, L.dst, args, L.n_tiles);
}

// Tile bits for source-low a and result-low b: source bits [0, a) plus the
// source bits of result bits [0, b).
int tile_bits(const int *inv, int a, int b) {
  int u = a;
  for (int p = 0; p < b; ++p) u += inv[p] >= a;
  return u;
}
bool tile_ok(const int *inv, int a, int b, int nb) {
  int u = tile_bits(inv, a, b);
  return a <= nb && b <= nb && u <= 12 && u - a <= 6 && u - b <= 6;
}
// Tile arguments of the bit permutation sigma (source bit -> result bit,
// inv its inverse) for the tile shape cfg.  False on an internal mismatch.
bool plan_bits(const int *sigma, const int *inv, int nb, const PermTune &cfg, PermBitsArgs &args, int *u_out,
               int *no_out) {
  memset(&args, 0, sizeof(args));
  int a = std::min(cfg.a, nb), b = std::min(cfg.b, nb);
  if (!tile_ok(inv, a, b, nb)) a = b = std::min(6, nb);
  const bool pad = (cfg.order & 2) != 0;
  // Grow the tile while it is smaller than the target (low bits kept in place,
  // or a low result bit taken from a high source bit, give small tiles).
  while (!pad && tile_bits(inv, a, b) < cfg.u_target) {
    int na = std::min(a + 1, nb), nbb = std::min(b + 1, nb);
    if ((na == a && nbb == b) || !tile_ok(inv, na, nbb, nb)) break;
    a = na;
    b = nbb;
  }
  // Tile bits: source bits [0, a), the source bits of result bits [0, b) not
  // already in, then (pad) outer bits by distance of their source and result
  // positions from 10.
  int extra[6], ne = 0;
  for (int p = 0; p < b; ++p)
    if (inv[p] >= a) extra[ne++] = inv[p];  // ordered by result position
  if (pad) {
    bool in_tile[64] = {};
    for (int p = 0; p < a; ++p) in_tile[p] = true;
    for (int m = 0; m < ne; ++m) in_tile[extra[m]] = true;
    auto dist = [&](int q) { return std::abs(q - 10) + std::abs(sigma[q] - 10); };
    while (a + ne < std::min(cfg.u_target, nb) && ne < 6 && a + ne - b < 6) {
      int best = -1;
      for (int q = 0; q < nb; ++q)
        if (!in_tile[q] && (best < 0 || dist(q) < dist(best))) best = q;
      if (best < 0) break;
      in_tile[best] = true;
      extra[ne++] = best;
    }
  }
  const int u = a + ne;
  // t position of each tile source bit
  int tpos[64];
  for (int p = 0; p < 64; ++p) tpos[p] = -1;
  for (int p = 0; p < a; ++p) tpos[p] = p;
  for (int m = 0; m < ne; ++m) tpos[extra[m]] = a + m;
  for (int j = 0; j < (1 << ne); ++j) {
    uint64_t off = 0;
    for (int m = 0; m < ne; ++m)
      if ((j >> m) & 1) off |= 1ull << extra[m];
    args.in_hi[j] = off;
  }
  // Result-order index s: bits [0, b) = result bits [0, b); above them the
  // remaining tile bits (source bits in [0, a) or padding whose result bit
  // >= b), by result position.
  int rest[6], nr = 0;
  for (int p = b; p < nb; ++p)
    if (tpos[inv[p]] >= 0) {
      if (nr == 6) return false;
      rest[nr++] = p;  // result positions
    }
  if (b + nr != u) return false;
  for (int j = 0; j < (1 << nr); ++j) {
    uint64_t off = 0;
    for (int m = 0; m < nr; ++m)
      if ((j >> m) & 1) off |= 1ull << rest[m];
    args.out_hi[j] = off;
  }
  int s_to_tbit[12];
  for (int p = 0; p < b; ++p) s_to_tbit[p] = tpos[inv[p]];
  for (int m = 0; m < nr; ++m) s_to_tbit[b + m] = tpos[inv[rest[m]]];
  for (int j = 0; j < 64; ++j) {
    uint32_t lo = 0, hi = 0;
    for (int m = 0; m < 6; ++m) {
      if (!((j >> m) & 1)) continue;
      if (m < u) lo |= 1u << s_to_tbit[m];
      if (m + 6 < u) hi |= 1u << s_to_tbit[m + 6];
    }
    args.s2t_lo[j] = (uint16_t)lo;
    args.s2t_hi[j] = (uint16_t)hi;
  }
  // Outer (tile index) bits, fastest first.
  int outer[64], no = 0;
  for (int p = 0; p < nb; ++p)
    if (tpos[p] < 0) outer[no++] = p;
  if (cfg.order & 4)
    std::stable_sort(outer, outer + no, [&](int x, int y) {
      return std::abs(x - 10) + std::abs(sigma[x] - 10) < std::abs(y - 10) + std::abs(sigma[y] - 10);
    });
  else if (cfg.order & 1)
    std::stable_sort(outer, outer + no,
                     [&](int x, int y) { return std::max(x, sigma[x]) < std::max(y, sigma[y]); });
  for (int j = 0; j < no; ++j) {
    args.ob_in[j] = (uint8_t)outer[j];
    args.ob_out[j] = (uint8_t)sigma[outer[j]];
  }
  args.a = a;
  args.b = b;
  args.u = u;
  args.n_outer = no;
  if (u > 12 || u < 0 || no + u != nb) return false;
  *u_out = u;
  *no_out = no;
  return true;
}

void launch_bits(hipStream_t stream, int num_cus, const c128 *src, c128 *dst, const PermBitsArgs &args, int u, int no,
                 int nb, const PermTune &cfg) {
  const uint64_t n_tiles = 1ull << no;
  const size_t lds = (size_t)sizeof(c128) << u;
  const uint64_t cap = (uint64_t)std::max(1, num_cus) * (uint64_t)cfg.blocks_per_cu;
  unsigned blocks = (unsigned)std::min<uint64_t>(std::min<uint64_t>(n_tiles, cap), 1u << 30);
  PermLaunch L{stream, blocks, lds, src, dst, n_tiles, nb <= 28, cfg.nt != 0};
  switch (u) {
    case 0: launch_tile<0>(L, args); break;
    case 1: launch_tile<1>(L, args); break;
    case 2: launch_tile<2>(L, args); break;
    case 3: launch_tile<3>(L, args); break;
    case 4: launch_tile<4>(L, args); break;
    case 5: launch_tile<5>(L, args); break;
    case 6: launch_tile<6>(L, args); break;
    case 7: launch_tile<7>(L, args); break;
    case 8: launch_tile<8>(L, args); break;
    case 9: launch_tile<9>(L, args); break;
    case 10: launch_tile<10>(L, args); break;
    case 11: launch_tile<11>(L, args); break;
    default: launch_tile<12>(L, args); break;
  }
}

// Times the candidate tile shapes for one signature on the caller's buffers and
// returns the fastest; *wrote is set when dst already holds the result.  Every
// shape is timed once, then the two fastest again (the better second timing
// wins), so one noisy timing does not decide.  launch(args, u, no, cfg) runs
// one shape and returns false when the kernel does not take it.
template <class Launch>
qtnh_status tune_shapes(qtnh_ctx_t ctx, const int *sigma, const int *inv, int nb, Launch launch, PermTune *best,
                        bool *wrote) {
  *wrote = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  // Time of kPermTimedRuns launches of shape cand (after one warm-up launch); < 0 if it does not apply.
  auto time_shape = [&](const PermTune &cand) -> float {
    PermBitsArgs args;
    int u, no;
    // A filled tile differs from a grown one only when the natural tile is short.
    if ((cand.order & 2) && tile_bits(inv, std::min(cand.a, nb), std::min(cand.b, nb)) >= std::min(cand.u_target, nb))
      return -1.0f;
    if (err != hipSuccess || !plan_bits(sigma, inv, nb, cand, args, &u, &no)) return -1.0f;
    if (!launch(args, u, no, cand)) return -1.0f;  // warm-up; writes the result
    err = hipEventRecord(e0, ctx->stream);
    for (int r = 0; r < kPermTimedRuns && err == hipSuccess; ++r) {
      launch(args, u, no, cand);
      err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipEventRecord(e1, ctx->stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err != hipSuccess) return -1.0f;
    *wrote = true;
    return ms;
  };
  constexpr int kCand = (int)(sizeof(kPermCandidates) / sizeof(kPermCandidates[0]));
  int first = -1, second = -1;
  float t_first = 0.0f, t_second = 0.0f;
  for (int i = 0; i < kCand; ++i) {
    const float ms = time_shape(kPermCandidates[i]);
    if (ms < 0.0f) continue;
    if (first < 0 || ms < t_first) {
      second = first;
      t_second = t_first;
      first = i;
      t_first = ms;
    } else if (second < 0 || ms < t_second) {
      second = i;
      t_second = ms;
    }
  }
  if (first >= 0) *best = kPermCandidates[first];
  if (first >= 0 && second >= 0) {
    const float a = time_shape(kPermCandidates[first]);
    const float b = time_shape(kPermCandidates[second]);
    if (a >= 0.0f && b >= 0.0f && b < a) *best = kPermCandidates[second];
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return check_hip(ctx, err, "permute tuning");
}

qtnh_status tune_bits(qtnh_ctx_t ctx, const c128 *src, c128 *dst, const int *sigma, const int *inv, int nb,
                      PermTune *best, bool *wrote) {
  auto launch = [&](const PermBitsArgs &args, int u, int no, const PermTune &cfg) {
    launch_bits(ctx->stream, ctx->num_cus, src, dst, args, u, no, nb, cfg);
    return true;
  };
  return tune_shapes(ctx, sigma, inv, nb, launch, best, wrote);
}

// The tuned tile shape cached for a signature key, or (not capturing, large
// enough) the fastest of the candidates timed with launch; else the default.
// *wrote: the timing runs already wrote the result.
template <class Launch>
qtnh_status tuned_shape(qtnh_ctx_t ctx, const std::string &key, uint64_t n, const int *sigma, const int *inv, int nb,
                        Launch launch, PermTune *cfg, bool *wrote) {
  *wrote = false;
  *cfg = g_perm;
  if (!g_perm_auto || n < kPermAutoMin) return QTNH_OK;
  {
    std::lock_guard<std::mutex> lock(g_perm_mu);
    auto it = g_perm_cache.find(key);
    if (it != g_perm_cache.end()) {
      *cfg = it->second;
      return QTNH_OK;
    }
  }
  if (ctx->fixed_shapes) return QTNH_OK;  // no timing (events, synchronisation) in this call
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess) {
    (void)hipGetLastError();  // not tuned: the query's error is not the caller's
    return QTNH_OK;
  }
  if (cap != hipStreamCaptureStatusNone) return QTNH_OK;
  qtnh_status st = tune_shapes(ctx, sigma, inv, nb, launch, cfg, wrote);
  if (st) return st;
  std::lock_guard<std::mutex> lock(g_perm_mu);
  if (g_perm_cache.size() >= kPermCacheMax) g_perm_cache.clear();
  g_perm_cache[key] = *cfg;
  return QTNH_OK;
}

template <int NB, int U>
void launch_gate_u(const PermLaunch &L, const PermBitsArgs &args, const GateRowArgs<NB> &gr) {
  if (L.narrow)
    hipLaunchKernelGGL((permute_gate_kernel<U, uint32_t, NB>), dim3(L.blocks), dim3(kPermThreads), L.lds, L.stream,
                       L.src, L.dst, args, gr, L.n_tiles);
  else
    hipLaunchKernelGGL((permute_gate_kernel<U, uint64_t, NB>), dim3(L.blocks), dim3(kPermThreads), L.lds, L.stream,
                       L.src, L.dst, args, gr, L.n_tiles);
}

// One launch of the fused gate kernel; false for tiles under 10 bits (the
// kernels are built for the 10- to 12-bit tiles the tuned shapes use).
template <int NB>
bool launch_gate_rows(hipStream_t stream, int num_cus, const c128 *src, c128 *dst, const PermBitsArgs &args, int u,
                      int no, int nb, const PermTune &cfg, const GateRowArgs<NB> &gr) {
  const uint64_t n_tiles = 1ull << no;
  const uint64_t cap = (uint64_t)std::max(1, num_cus) * (uint64_t)cfg.blocks_per_cu;
  const unsigned blocks = (unsigned)std::min<uint64_t>(std::min<uint64_t>(n_tiles, cap), 1u << 30);
  PermLaunch L{stream, blocks, (size_t)sizeof(c128) << u, src, dst, n_tiles, nb <= 28, true};
  switch (u) {
    case 10: launch_gate_u<NB, 10>(L, args, gr); return true;
    case 11: launch_gate_u<NB, 11>(L, args, gr); return true;
    case 12: launch_gate_u<NB, 12>(L, args, gr); return true;
    default: return false;
  }
}

template <int NB>
qtnh_status permute_gate(qtnh_ctx_t ctx, const c128 *A, c128 *C, const c128 *B, int nb, const int *sigma,
                         const uint32_t *xb, bool *done) {
  *done = false;
  int inv[64];
  for (int p = 0; p < nb; ++p) inv[sigma[p]] = p;
  GateRowArgs<NB> gr;
  gr.B = B;
  for (int i = 0; i < (1 << NB) * (1 << NB); ++i) gr.xb[i] = xb[i];
  auto launch = [&](const PermBitsArgs &args, int u, int no, const PermTune &cfg) {
    return launch_gate_rows<NB>(ctx->stream, ctx->num_cus, A, C, args, u, no, nb, cfg, gr);
  };
  std::string key((const char *)&ctx->device, sizeof(int));
  key.push_back('G');
  key.push_back((char)NB);
  for (int p = 0; p < nb; ++p) key.push_back((char)sigma[p]);
  PermTune cfg;
  bool wrote = false;
  qtnh_status st = tuned_shape(ctx, key, 1ull << nb, sigma, inv, nb, launch, &cfg, &wrote);
  if (st) return st;
  if (wrote) {
    *done = true;
    return QTNH_OK;
  }
  PermBitsArgs args;
  int u, no;
  if (!plan_bits(sigma, inv, nb, cfg, args, &u, &no) || !launch(args, u, no, cfg)) return QTNH_OK;
  *done = true;
  return check_hip(ctx, hipGetLastError(), "permute_gate launch");
}
}  // namespace

qtnh_status launch_permute_gate(qtnh_ctx_t ctx, const c128 *A, c128 *C, const c128 *B, int nb, const int *sigma,
                                int gate_bits, const uint32_t *xb, bool *done) {
  *done = false;
  if (nb < 12 || nb > 64) return QTNH_OK;
  if (gate_bits == 1) return permute_gate<1>(ctx, A, C, B, nb, sigma, xb, done);
  if (gate_bits == 2) return permute_gate<2>(ctx, A, C, B, nb, sigma, xb, done);
  if (gate_bits == 3) return permute_gate<3>(ctx, A, C, B, nb, sigma, xb, done);
  return QTNH_OK;
}

qtnh_status launch_permute(qtnh_ctx_t ctx, const c128 *src, c128 *dst, int rank, const uint64_t *dims,
                           const uint16_t *ptup) {
  uint64_t n;
  checked_size(rank, dims, &n);
  bool identity = true;
  for (int i = 0; i < rank; ++i) identity = identity && (ptup[i] == i);
  // Result dims and strides.
  uint64_t nd[kMaxRank];
  for (int i = 0; i < rank; ++i) nd[ptup[i]] = dims[i];
  if (identity) return launch_copy(ctx, src, dst, n);

  BitLayout in = bit_layout(rank, dims);
  BitLayout out = bit_layout(rank, nd);
  if (in.binary && out.binary) {
    const int nb = in.nbits;
    // sigma: source bit -> result bit.
    int sigma[64], inv[64];
    for (int i = 0; i < rank; ++i)
      for (int t = 0; t < in.width[i]; ++t) sigma[in.shift[i] + t] = out.shift[ptup[i]] + t;
    for (int p = 0; p < nb; ++p) inv[sigma[p]] = p;
    PermTune cfg = g_perm;
    if (g_perm_auto && n >= kPermAutoMin) {
      std::string key((const char *)&ctx->device, sizeof(int));
      for (int p = 0; p < nb; ++p) key.push_back((char)sigma[p]);
      bool found = false;
      {
        std::lock_guard<std::mutex> lock(g_perm_mu);
        auto it = g_perm_cache.find(key);
        if (it != g_perm_cache.end()) {
          cfg = it->second;
          found = true;
        }
      }
      hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
      if (!found && ctx->fixed_shapes) found = true;  // keep the default shape: no timing in this call
      hipError_t qe = found ? hipSuccess : hipStreamIsCapturing(ctx->stream, &cap);
      if (qe != hipSuccess) (void)hipGetLastError();  // not tuned: the query's error is not the caller's
      if (!found && qe == hipSuccess && cap == hipStreamCaptureStatusNone) {
        bool wrote = false;
        qtnh_status st = tune_bits(ctx, src, dst, sigma, inv, nb, &cfg, &wrote);
        if (st) return st;
        std::lock_guard<std::mutex> lock(g_perm_mu);
        if (g_perm_cache.size() >= kPermCacheMax) g_perm_cache.clear();
        g_perm_cache[key] = cfg;
        if (wrote) return QTNH_OK;
      }
    }
    PermBitsArgs args;
    int u, no;
    if (!plan_bits(sigma, inv, nb, cfg, args, &u, &no))
      return set_error(ctx, QTNH_EINVAL, "permute: internal tile mismatch");
    launch_bits(ctx->stream, ctx->num_cus, src, dst, args, u, no, nb, cfg);
    return check_hip(ctx, hipGetLastError(), "permute_bits launch");
  }

  PermGenArgs g;
  memset(&g, 0, sizeof(g));
  uint64_t st[kMaxRank];
  uint64_t s = 1;
  for (int i = rank - 1; i >= 0; --i) {
    st[i] = s;
    s *= dims[i];
  }
  for (int i = 0; i < rank; ++i) {
    g.nd[ptup[i]] = dims[i];
    g.jump[ptup[i]] = st[i];
  }
  g.rank = rank;
  unsigned blocks = grid_for(n, 256, 65536);
  hipLaunchKernelGGL(permute_general_kernel, dim3(blocks), dim3(256), 0, ctx->stream, src, dst, g, n);
  return check_hip(ctx, hipGetLastError(), "permute_general launch");
}

qtnh_status launch_swap(qtnh_ctx_t ctx, c128 *buf, int rank, const uint64_t *dims, int i, int j) {
  uint64_t n;
  checked_size(rank, dims, &n);
  BitLayout L = bit_layout(rank, dims);
  if (L.binary && dims[i] == 2) {
    // axis i < j: axis j holds the lower bit
    int lo = L.shift[j], hi = L.shift[i];
    uint64_t pairs = n >> 2;
    unsigned blocks = grid_for(pairs, 256, 65536);
    hipLaunchKernelGGL(swap_bits_kernel, dim3(blocks), dim3(256), 0, ctx->stream, buf, lo, hi, pairs);
    return check_hip(ctx, hipGetLastError(), "swap_bits launch");
  }
  uint64_t si = 1, sj = 1;
  for (int k = rank - 1; k > i; --k) si *= dims[k];
  for (int k = rank - 1; k > j; --k) sj *= dims[k];
  unsigned blocks = grid_for(n, 256, 65536);
  hipLaunchKernelGGL(swap_general_kernel, dim3(blocks), dim3(256), 0, ctx->stream, buf, n, si, sj, dims[i]);
  return check_hip(ctx, hipGetLastError(), "swap_general launch");
}


namespace {
// Many bit permutations in one launch (the small operand permutes of a
// network batch's steps): block b of descriptor D moves dst elements
// [1024 b, 1024 b + 1024) of D, thread t the elements 1024 b + t + 256 i,
// so each store instruction writes 4 KiB of consecutive elements.  The source
// offset of an element splits over its disjoint bits: that of the block's
// base (bits 10 and up) OR that of t + 256 i (bits 0..9).
__global__ void __launch_bounds__(256) permute_grouped_kernel(const PermDesc *__restrict__ d, int nd,
                                                              uint64_t total) {
  for (uint64_t blk = blockIdx.x; blk < total; blk += gridDim.x) {
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (d[mid].first <= blk)
        lo = mid;
      else
        hi = mid - 1;
    }
    const PermDesc &D = d[lo];
    const uint64_t base = (blk - D.first) << 10;
    uint64_t sbase = 0;
    for (int p = 10; p < D.nb; ++p)
      if ((base >> p) & 1ull) sbase |= 1ull << D.sbit[p];
    const int lowbits = D.nb < 10 ? D.nb : 10;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t low = threadIdx.x + 256u * i;
      const uint64_t j = base + low;
      if (j >= D.n) continue;
      uint64_t s = sbase;
      for (int p = 0; p < lowbits; ++p)
        if ((low >> p) & 1u) s |= 1ull << D.sbit[p];
      D.dst[j] = D.src[s];
    }
  }
}
}  // namespace

bool permute_source_bits(int rank, const uint64_t *dims, const uint16_t *ptup, int *nb, int8_t *sbit) {
  uint64_t nd[kMaxRank];
  for (int i = 0; i < rank; ++i) nd[ptup[i]] = dims[i];
  const BitLayout in = bit_layout(rank, dims);
  const BitLayout out = bit_layout(rank, nd);
  if (!in.binary || !out.binary || in.nbits > 64) return false;
  for (int i = 0; i < rank; ++i)
    for (int t = 0; t < in.width[i]; ++t) sbit[out.shift[ptup[i]] + t] = (int8_t)(in.shift[i] + t);
  *nb = in.nbits;
  return true;
}

qtnh_status launch_permute_grouped(qtnh_ctx_t ctx, const PermDesc *d, int nd, uint64_t total) {
  if (nd <= 0 || total == 0) return QTNH_OK;
  const unsigned grid = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus * 16);
  hipLaunchKernelGGL(permute_grouped_kernel, dim3(grid), dim3(256), 0, ctx->stream, d, nd, total);
  return check_hip(ctx, hipGetLastError(), "grouped permute launch");
}

}  // namespace qtnh_hip
