// In-place symmetric-gate and diagonal-gate application.
//
// Gate: psi'[.. o@axes ..] = sum_i psi[.. i@axes ..] * G[i, o]
// (reference src/tensor/contract.cpp:250-289 placement, :131-162 MAC loop).
// The sum runs over i in row-major order starting from 0, each complex
// product rounded as (ar*br - ai*bi, ar*bi + ai*br) without FMA (this file is
// built with -ffp-contract=off): the same bits as the reference's
// std::complex<double> loop.
//
// Diagonal: psi[x] *= d[x@axes] (intended DiagTensor, reference diag.hpp:11-14).
//
// Power-of-two dims with at most 3 gate bits run the streaming kernels below:
// one thread per group of 2^kb elements that differ only in the gate bits.
// Anything else goes through the general contraction into scratch memory.
#include <algorithm>
#include <cstring>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

constexpr int kGateThreads = 256;

// Most workgroups of the streaming gate and diagonal kernels (they loop over
// the rest): enough to keep many requests in flight on every CU.
unsigned g_stream_blocks = 1u << 20;
// Grid limit of the kernels sized to one element (or row) per thread.
constexpr unsigned kMaxStreamBlocks = 1u << 30;

__device__ inline c128 cmul(c128 a, c128 b) {
  c128 r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}

template <int KB>
struct GateArgs {
  c128 g[(1 << KB) * (1 << KB)];  // G[i][o] at g[i * 2^KB + o]; 3-qubit gates: at g[8 * o + i]
  uint64_t comb[1 << KB];         // offset of gate combination c from the group base
  int zpos[KB];                   // gate bit positions, ascending (zeros inserted here)
};

typedef double double2_t __attribute__((ext_vector_type(2)));

// U groups per thread and iteration, all loaded before any is stored: more
// memory requests in flight per thread.  NT: non-temporal loads and stores,
// used when the lowest gate bit is >= 3, so that the 64 lanes of each load
// cover whole 128-byte lines (with lower gate bits the lanes of one load are
// strided inside lines, and the next load of the same lines must hit in cache).
//
// 3-qubit gates: the 128 doubles of the coefficients do not fit the scalar
// registers, so each output's 8 (one column of G, contiguous in the argument
// block) are loaded by scalar loads where they are used; the output index is
// hidden from the optimiser, which otherwise hoists all 64 out of the loop and
// spills them into vector registers.  An LDS copy read per use measured
// slower.
template <int KB, bool NT>
__global__ void __launch_bounds__(kGateThreads) gate_bits_kernel(c128 *__restrict__ psi, GateArgs<KB> args,
                                                                 uint64_t n_groups) {
  constexpr int K = 1 << KB;
  constexpr int U = KB <= 2 ? 2 : 1;
  double2_t *p2 = reinterpret_cast<double2_t *>(psi);
  const uint64_t stride = (uint64_t)gridDim.x * kGateThreads;
  for (uint64_t g0 = (uint64_t)blockIdx.x * kGateThreads + threadIdx.x; g0 < n_groups; g0 += stride * U) {
    uint64_t base[U];
    double2_t in[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t g = g0 + u * stride;
      uint64_t b = g < n_groups ? g : g0;
#pragma unroll
      for (int j = 0; j < KB; ++j) {
        const int p = args.zpos[j];
        b = ((b >> p) << (p + 1)) | (b & ((1ull << p) - 1));
      }
      base[u] = b;
#pragma unroll
      for (int c = 0; c < K; ++c) in[u][c] = NT ? __builtin_nontemporal_load(p2 + b + args.comb[c]) : p2[b + args.comb[c]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (g0 + u * stride >= n_groups) continue;
#pragma unroll
      for (int o = 0; o < K; ++o) {
        int oc = o;
        if (KB == 3) asm volatile("" : "+s"(oc));
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const c128 x = {in[u][i].x, in[u][i].y};
          const c128 gg = KB == 3 ? args.g[K * oc + i] : args.g[i * K + o];
          const c128 pr = cmul(x, gg);
          ar = ar + pr.re;
          ai = ai + pr.im;
        }
        if (NT)
          __builtin_nontemporal_store(double2_t{ar, ai}, p2 + base[u] + args.comb[o]);
        else
          p2[base[u] + args.comb[o]] = double2_t{ar, ai};
      }
    }
  }
}

constexpr int kMaxDiag = 64;

struct DiagArgs {
  c128 d[kMaxDiag];
  int bitpos[6];  // bit position of diagonal-index bit (kb-1-j), i.e. MSB first
  int kb;
};

// Tables of up to 4 entries: the entries are read once per thread with
// uniform (scalar) loads and chosen per element by selects in registers.  An
// index into the argument block that varies per lane compiles to a vector
// memory load per element, which queues behind the stream's own loads (a
// diagonal on the two lowest bits measured 1.07 TB/s that way on one box).
__global__ void __launch_bounds__(kGateThreads) diag_bits_kernel(c128 *__restrict__ psi, DiagArgs args, uint64_t n) {
  constexpr int U = 4;
  const c128 t0 = args.d[0], t1 = args.d[1], t2 = args.d[2], t3 = args.d[3];
  double2_t *p2 = reinterpret_cast<double2_t *>(psi);
  const uint64_t stride = (uint64_t)gridDim.x * kGateThreads;
  for (uint64_t e0 = (uint64_t)blockIdx.x * kGateThreads + threadIdx.x; e0 < n; e0 += stride * U) {
    double2_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t e = e0 + u * stride;
      v[u] = e < n ? __builtin_nontemporal_load(p2 + e) : double2_t{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t e = e0 + u * stride;
      if (e >= n) continue;
      uint32_t c = 0;
      for (int j = 0; j < args.kb; ++j) c = (c << 1) | (uint32_t)((e >> args.bitpos[j]) & 1ull);
      const c128 lo = (c & 1) ? t1 : t0, hi = (c & 1) ? t3 : t2;
      const c128 r = cmul(c128{v[u].x, v[u].y}, (c & 2) ? hi : lo);
      __builtin_nontemporal_store(double2_t{r.re, r.im}, p2 + e);
    }
  }
}

// As diag_bits_kernel, for tables of more than 4 entries: the table is copied
// into LDS once per workgroup and read from there per element (indexing the
// argument block per lane would issue a memory load per element, which waits
// behind the stream's own loads).  Same arithmetic as diag_bits_kernel.
__global__ void __launch_bounds__(kGateThreads) diag_bits_lds_kernel(c128 *__restrict__ psi, DiagArgs args,
                                                                     uint64_t n) {
  constexpr int U = 4;
  __shared__ double2_t tab[kMaxDiag];
  const int K = 1 << args.kb;
  for (int i = threadIdx.x; i < K; i += kGateThreads) tab[i] = double2_t{args.d[i].re, args.d[i].im};
  __syncthreads();
  double2_t *p2 = reinterpret_cast<double2_t *>(psi);
  const uint64_t stride = (uint64_t)gridDim.x * kGateThreads;
  for (uint64_t e0 = (uint64_t)blockIdx.x * kGateThreads + threadIdx.x; e0 < n; e0 += stride * U) {
    double2_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t e = e0 + u * stride;
      v[u] = e < n ? __builtin_nontemporal_load(p2 + e) : double2_t{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t e = e0 + u * stride;
      if (e >= n) continue;
      uint32_t c = 0;
      for (int j = 0; j < args.kb; ++j) c = (c << 1) | (uint32_t)((e >> args.bitpos[j]) & 1ull);
      const double2_t dc = tab[c];
      const c128 r = cmul(c128{v[u].x, v[u].y}, c128{dc.x, dc.y});
      __builtin_nontemporal_store(double2_t{r.re, r.im}, p2 + e);
    }
  }
}

struct DiagGenArgs {
  c128 d[kMaxDiag];
  uint64_t stride[kMaxRank];  // state stride of gate axis m
  uint64_t dim[kMaxRank];     // dim of gate axis m
  int k;
};

__global__ void diag_general_kernel(c128 *__restrict__ psi, DiagGenArgs args, uint64_t n) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t c = 0;
    for (int m = 0; m < args.k; ++m) c = c * args.dim[m] + (e / args.stride[m]) % args.dim[m];
    psi[e] = cmul(psi[e], args.d[c]);
  }
}

// Gates with bits among the lowest 6 ("lane bits": the element's position in
// a wave's 64 consecutive elements).  Thread t handles the elements whose
// non-gate bits and lane gate bits come from t (consecutive threads,
// consecutive elements) and loads one element per combination of the high
// gate bits: every load and store of a wave covers 1 KiB contiguous, streamed
// non-temporally.  The inputs held by partner lanes (other lane-bit values)
// arrive by cross-lane shuffles; each lane computes only its own outputs, with
// the same sum order and roundings as gate_bits_kernel.
template <int KB, int KH>
struct LaneGateArgs {
  c128 g[(1 << KB) * (1 << KB)];  // G[i][o]
  uint64_t hoff[1 << KH];         // element offset of high-bit combination h
  int hzpos[KH > 0 ? KH : 1];     // high gate bits, ascending (zeros inserted here)
  uint32_t lane_mask;             // lane gate bits
  uint8_t src_lbits[1 << KB];     // input i: its lane gate bits
  uint8_t src_h[1 << KB];         // input i: its high combination
  int8_t jpos[KB];                // gate index bit KB-1-j: state bit
  int8_t jh[KB];                  // ... its bit in the high combination (-1: a lane bit)
};

template <int KB, int KH>
__global__ void __launch_bounds__(kGateThreads) gate_lanes_kernel(c128 *__restrict__ psi, LaneGateArgs<KB, KH> args,
                                                                  uint64_t n_rows) {
  constexpr int K = 1 << KB, H = 1 << KH;
  // Rows per thread per step: their loads are all issued before the first
  // shuffle, so a wave keeps several lines in flight.
  constexpr int U = KH >= 2 ? 1 : 4 >> KH;
  double2_t *p2 = reinterpret_cast<double2_t *>(psi);
  const int lane = threadIdx.x & 63;
  // The output index o of high combination h depends only on the lane, so the
  // coefficients G[i][o] a thread uses are fixed: they are read once instead of
  // per element from the argument block (a per-lane vector-memory load whose
  // wait would also wait for the stream's loads and stores in flight).  Held in
  // registers when all gate bits are lane bits (measured faster there: the
  // exchanges already load the LDS pipe), in LDS otherwise (faster with high
  // gate bits, e.g. 77% -> 79% of HBM peak for a 2-qubit gate with one).
  constexpr bool kRegCoef = KH == 0;
  int outs[H];
#pragma unroll
  for (int h = 0; h < H; ++h) {
    int o = 0;
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      const int bit = args.jh[j] >= 0 ? ((h >> args.jh[j]) & 1) : ((lane >> args.jpos[j]) & 1);
      o |= bit << (KB - 1 - j);
    }
    outs[h] = o;
  }
  double2_t greg[kRegCoef ? H : 1][kRegCoef ? K : 1];
  __shared__ double2_t gco[kRegCoef ? 1 : K * K];
  if constexpr (kRegCoef) {
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
      for (int i = 0; i < K; ++i) greg[h][i] = double2_t{args.g[i * K + outs[h]].re, args.g[i * K + outs[h]].im};
  } else {
    for (int i = threadIdx.x; i < K * K; i += kGateThreads) gco[i] = double2_t{args.g[i].re, args.g[i].im};
    __syncthreads();
  }
  const uint64_t stride = (uint64_t)gridDim.x * kGateThreads;
  // Every lane of a wave takes part in the shuffles: the loop bounds are
  // wave-uniform (n_rows is a multiple of 64 and so is the stride).
  for (uint64_t t0 = (uint64_t)blockIdx.x * kGateThreads + threadIdx.x; t0 < n_rows; t0 += stride * U) {
    // LDS coefficients are read at each use: without this clobber the
    // compiler hoists the loop-invariant reads into registers (218 VGPRs for
    // a 3-qubit gate with two high bits, 2 waves per SIMD; 69 with it, and
    // 4.6 -> 6.2 TB/s at 2^28 elements).
    if constexpr (!kRegCoef) asm volatile("" ::: "memory");
    double2_t x[U][H];
    uint64_t e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t t = t0 + (uint64_t)u * stride;
      if (t >= n_rows) break;  // wave-uniform
      e[u] = t;
#pragma unroll
      for (int j = 0; j < KH; ++j) {
        const int p = args.hzpos[j];
        e[u] = ((e[u] >> p) << (p + 1)) | (e[u] & ((1ull << p) - 1));
      }
#pragma unroll
      for (int h = 0; h < H; ++h) x[u][h] = __builtin_nontemporal_load(p2 + e[u] + args.hoff[h]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (t0 + (uint64_t)u * stride >= n_rows) break;  // wave-uniform
      // All K inputs of this lane's groups: input i sits in lane (lane & ~mask) | lbits(i), slot h(i).
      double2_t in[K];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int src = (lane & ~(int)args.lane_mask) | args.src_lbits[i];
        const double2_t v = x[u][args.src_h[i]];
        in[i].x = __shfl(v.x, src, 64);
        in[i].y = __shfl(v.y, src, 64);
      }
#pragma unroll
      for (int h = 0; h < H; ++h) {
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const c128 xi = {in[i].x, in[i].y};
          const double2_t gi = kRegCoef ? greg[kRegCoef ? h : 0][kRegCoef ? i : 0] : gco[i * K + outs[h]];
          const c128 pr = cmul(xi, c128{gi.x, gi.y});
          ar = ar + pr.re;
          ai = ai + pr.im;
        }
        __builtin_nontemporal_store(double2_t{ar, ai}, p2 + e[u] + args.hoff[h]);
      }
    }
  }
}

// 3-qubit gates whose bits are all among the lowest 6.  Exchanging lanes
// would take 16 cross-lane moves per element; instead each wave stages 8 rows
// of 64 elements (a 512-element tile, 8 KiB) in its own LDS region with
// coalesced loads, each lane then takes one of the tile's 64 groups of 8
// elements and applies the gate as gate_bits_kernel does (same arithmetic,
// same order), and the results leave through LDS with coalesced stores.  The
// coefficients are wave-uniform: each output's 8 come from the argument block
// by scalar loads (an LDS copy read per use measured 1.5% slower: its 64
// 16-byte broadcast reads per tile and lane doubled the kernel's LDS traffic).  The
// LDS slot of tile element e is e ^ ((e / 8) % 8): conflict-free for the row
// accesses and for group reads when the gate bits are the lowest three (at
// most 2-way otherwise).  Waves synchronise only with themselves (their
// region is private), so every wave loops over its own tiles.
struct Low3Args {
  c128 gt[64];       // G[i][o] at gt[8 * o + i]: each output's coefficients contiguous
  uint16_t coff[8];  // tile offset of gate combination c (MSB-first gate bits)
  uint16_t gbase[64];  // tile offset of group q: its bits spread over the non-gate tile bits
};

__device__ inline uint32_t low3_slot(uint32_t e) { return e ^ ((e >> 3) & 7u); }

__global__ void __launch_bounds__(kGateThreads) gate_low3_kernel(c128 *__restrict__ psi, Low3Args args,
                                                                 uint64_t n_tiles) {
  constexpr int kWaves = kGateThreads / 64;
  __shared__ double2_t buf[kWaves][512];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double2_t *p2 = reinterpret_cast<double2_t *>(psi);
  double2_t *b = buf[w];
  const uint32_t gb = args.gbase[lane];
  const uint64_t step = (uint64_t)gridDim.x * kWaves;
  uint64_t t = (uint64_t)blockIdx.x * kWaves + w;
  if (t >= n_tiles) return;  // per wave: each wave synchronises only with itself
  double2_t v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = __builtin_nontemporal_load(p2 + (t << 9) + 64 * j + lane);
  for (;;) {
    double2_t *src = p2 + (t << 9);
#pragma unroll
    for (int j = 0; j < 8; ++j) b[low3_slot(64 * j + lane)] = v[j];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // The wave's next tile is loaded while this one is formed and stored.
    const uint64_t tn = t + step;
    const bool more = tn < n_tiles;  // wave-uniform
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = __builtin_nontemporal_load(p2 + (tn << 9) + 64 * j + lane);
    }
    double2_t in[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) in[c] = b[low3_slot(gb + args.coff[c])];
    // A lane's results go to its own group's slots, which no other lane
    // reads: each output is stored as soon as it is formed.
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      // The output index is hidden from the optimiser so that each output's
      // 8 coefficients are loaded (scalar loads) where they are used rather
      // than all 64 hoisted out of the loop.
      int oc = o;
      asm volatile("" : "+s"(oc));
      const c128 *col = args.gt + 8 * oc;
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const c128 pr = cmul(c128{in[i].x, in[i].y}, col[i]);
        ar = ar + pr.re;
        ai = ai + pr.im;
      }
      b[low3_slot(gb + args.coff[o])] = double2_t{ar, ai};
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int j = 0; j < 8; ++j) __builtin_nontemporal_store(b[low3_slot(64 * j + lane)], src + 64 * j + lane);
    if (!more) break;
    // The next tile's stores into b come after these reads.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    t = tn;
  }
}

qtnh_status launch_gate_low3(qtnh_ctx_t ctx, c128 *psi, uint64_t n, const int *msb_bits, const double *G) {
  Low3Args args;
  memset(&args, 0, sizeof(args));
  for (int i = 0; i < 8; ++i)
    for (int o = 0; o < 8; ++o) memcpy(&args.gt[8 * o + i], G + 2 * (8 * i + o), sizeof(c128));
  uint32_t gmask = 0;
  for (int j = 0; j < 3; ++j) gmask |= 1u << msb_bits[j];
  for (int c = 0; c < 8; ++c) {
    uint32_t off = 0;
    for (int j = 0; j < 3; ++j)
      if ((c >> (2 - j)) & 1) off |= 1u << msb_bits[j];
    args.coff[c] = (uint16_t)off;
  }
  // Group q: its 6 bits in order over the tile bits 0..8 that are not gate bits.
  for (uint32_t q = 0; q < 64; ++q) {
    uint32_t off = 0;
    int k = 0;
    for (int bit = 0; bit < 9; ++bit) {
      if ((gmask >> bit) & 1) continue;
      if ((q >> k) & 1) off |= 1u << bit;
      ++k;
    }
    args.gbase[q] = (uint16_t)off;
  }
  // One tile per wave (64 threads per tile), up to the streaming grid limit.
  // Rank 28: 5.52 TB/s.  Fewer waves each looping over tiles, with the next
  // tile's loads in flight, are slower: 5.38 at two tiles per wave, 5.29 at
  // sixteen, 5.09 (before the scalar coefficient loads) at 512 workgroups.
  const uint64_t n_tiles = n >> 9;
  const unsigned blocks = grid_for(n_tiles * 64, kGateThreads, g_stream_blocks);
  hipLaunchKernelGGL(gate_low3_kernel, dim3(blocks), dim3(kGateThreads), 0, ctx->stream, psi, args, n_tiles);
  return check_hip(ctx, hipGetLastError(), "gate launch");
}

template <int KB, int KH>
qtnh_status launch_gate_lanes(qtnh_ctx_t ctx, c128 *psi, uint64_t n, const int *msb_bits, const double *G) {
  constexpr int K = 1 << KB;
  LaneGateArgs<KB, KH> args;
  memset(&args, 0, sizeof(args));
  memcpy(args.g, G, sizeof(c128) * K * K);
  int hbits[KB], nh = 0;
  for (int j = 0; j < KB; ++j) {
    args.jpos[j] = (int8_t)msb_bits[j];
    if (msb_bits[j] < 6) {
      args.lane_mask |= 1u << msb_bits[j];
      args.jh[j] = -1;
    } else {
      hbits[nh] = msb_bits[j];
      args.jh[j] = (int8_t)nh++;
    }
  }
  for (int h = 0; h < (1 << KH); ++h) {
    uint64_t off = 0;
    for (int q = 0; q < KH; ++q)
      if ((h >> q) & 1) off |= 1ull << hbits[q];
    args.hoff[h] = off;
  }
  int z[KH > 0 ? KH : 1];
  for (int q = 0; q < KH; ++q) z[q] = hbits[q];
  std::sort(z, z + KH);
  for (int q = 0; q < KH; ++q) args.hzpos[q] = z[q];
  for (int i = 0; i < K; ++i) {
    uint32_t lb = 0;
    int h = 0;
    for (int j = 0; j < KB; ++j) {
      if (!((i >> (KB - 1 - j)) & 1)) continue;
      if (args.jh[j] < 0)
        lb |= 1u << msb_bits[j];
      else
        h |= 1 << args.jh[j];
    }
    args.src_lbits[i] = (uint8_t)lb;
    args.src_h[i] = (uint8_t)h;
  }
  // Grid (measured at 2^28 and 2^30 elements): most 1- and 2-qubit gates do
  // best with one row per thread, so that the workgroups in flight cover one
  // contiguous stretch of the state; 3-qubit gates need each thread's U rows
  // in flight together (one pass).
  const uint64_t rows = n >> KH;
  // Two rows per thread for 2-qubit gates on two lane bits (5.27 -> 5.63 TB/s
  // at 2^28 elements; one row was faster for the other 1- and 2-qubit shapes).
  constexpr uint64_t kRowsPerThread = KB == 2 && KH == 0 ? 2 : KB <= 2 ? 1 : (KH >= 2 ? 1 : 4 >> KH);
  unsigned blocks = grid_for((rows + kRowsPerThread - 1) / kRowsPerThread, kGateThreads, kMaxStreamBlocks);
  hipLaunchKernelGGL((gate_lanes_kernel<KB, KH>), dim3(blocks), dim3(kGateThreads), 0, ctx->stream, psi, args, rows);
  return check_hip(ctx, hipGetLastError(), "gate launch");
}

template <int KB>
qtnh_status launch_gate_bits(qtnh_ctx_t ctx, c128 *psi, uint64_t n, const int *msb_bits, const double *G) {
  constexpr int K = 1 << KB;
  GateArgs<KB> args;
  memset(&args, 0, sizeof(args));
  for (int i = 0; i < K; ++i)
    for (int o = 0; o < K; ++o) memcpy(&args.g[KB == 3 ? K * o + i : K * i + o], G + 2 * (K * i + o), sizeof(c128));
  for (int c = 0; c < K; ++c) {
    uint64_t off = 0;
    for (int j = 0; j < KB; ++j)
      if ((c >> (KB - 1 - j)) & 1) off |= 1ull << msb_bits[j];
    args.comb[c] = off;
  }
  int z[KB];
  for (int j = 0; j < KB; ++j) z[j] = msb_bits[j];
  std::sort(z, z + KB);
  for (int j = 0; j < KB; ++j) args.zpos[j] = z[j];
  // A gate bit among the lowest 6 of a large state: the lane-exchange kernel.
  int nlow = 0;
  for (int j = 0; j < KB; ++j) nlow += msb_bits[j] < 6;
  if (nlow > 0 && n >= 4096) {
    const int kh = KB - nlow;
    if (kh == 0 && KB == 3) return launch_gate_low3(ctx, psi, n, msb_bits, G);
    if (kh == 0) return launch_gate_lanes<KB, 0>(ctx, psi, n, msb_bits, G);
    if (kh == 1) return launch_gate_lanes<KB, (KB >= 2 ? 1 : 0)>(ctx, psi, n, msb_bits, G);
    return launch_gate_lanes<KB, (KB >= 3 ? 2 : 0)>(ctx, psi, n, msb_bits, G);
  }
  const uint64_t groups = n >> KB;
  constexpr int U = KB <= 2 ? 2 : 1;
  unsigned blocks = grid_for((groups + U - 1) / U, kGateThreads, g_stream_blocks);
  if (args.zpos[0] >= 3)
    hipLaunchKernelGGL((gate_bits_kernel<KB, true>), dim3(blocks), dim3(kGateThreads), 0, ctx->stream, psi, args, groups);
  else
    hipLaunchKernelGGL((gate_bits_kernel<KB, false>), dim3(blocks), dim3(kGateThreads), 0, ctx->stream, psi, args,
                       groups);
  return check_hip(ctx, hipGetLastError(), "gate launch");
}

}  // namespace

extern "C" void qtnh_tune_stream(int max_blocks) {
  g_stream_blocks = (unsigned)std::max(256, std::min(max_blocks, 1 << 22));
}

qtnh_status launch_gate(qtnh_ctx_t ctx, c128 *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                        const double *G) {
  uint64_t n;
  checked_size(rank, dims, &n);
  BitLayout L = bit_layout(rank, dims);
  int kb = 0;
  int msb_bits[64];
  if (L.binary) {
    for (int m = 0; m < k; ++m)
      for (int t = L.width[axes[m]] - 1; t >= 0; --t) msb_bits[kb++] = L.shift[axes[m]] + t;
  }
  if (L.binary && kb >= 1 && kb <= 3) {
    switch (kb) {
      case 1: return launch_gate_bits<1>(ctx, state, n, msb_bits, G);
      case 2: return launch_gate_bits<2>(ctx, state, n, msb_bits, G);
      default: return launch_gate_bits<3>(ctx, state, n, msb_bits, G);
    }
  }
  if (L.binary && kb == 0) return QTNH_OK;  // all gate axes of dim 1: G is 1x1
  // General case: contract into scratch under the symmetric placement, copy back.
  uint64_t K = 1;
  for (int m = 0; m < k; ++m) K *= dims[axes[m]];
  uint64_t gd[2 * kMaxRank];
  uint16_t wires[2 * kMaxRank], r1[kMaxRank], r2[2 * kMaxRank];
  for (int m = 0; m < k; ++m) {
    gd[m] = gd[m + k] = dims[axes[m]];
    wires[2 * m] = axes[m];
    wires[2 * m + 1] = (uint16_t)m;
  }
  int applies = 0;
  qtnh_status st = symm_repls(rank, 0, 2 * k, 0, k, wires, r1, r2, &applies);
  if (st || !applies) return set_error(ctx, QTNH_EINVAL, "gate: placement failed");
  void *ws;
  const uint64_t gbytes = K * K * sizeof(c128);
  const uint64_t gpad = (gbytes + 255) & ~255ull;
  st = workspace(ctx, gpad + n * sizeof(c128), &ws);
  if (st) return st;
  c128 *gdev = (c128 *)ws;
  c128 *tmp = (c128 *)((char *)ws + gpad);
  st = check_hip(ctx, hipMemcpyAsync(gdev, G, gbytes, hipMemcpyHostToDevice, ctx->stream), "gate H2D");
  if (st) return st;
  st = launch_contract_direct(ctx, state, rank, dims, gdev, 2 * k, gd, k, wires, r1, r2, tmp);
  if (st) return st;
  return check_hip(ctx, hipMemcpyAsync(state, tmp, n * sizeof(c128), hipMemcpyDeviceToDevice, ctx->stream),
                   "gate copy back");
}

qtnh_status launch_diag(qtnh_ctx_t ctx, c128 *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                        const double *d) {
  uint64_t n;
  checked_size(rank, dims, &n);
  uint64_t K = 1;
  for (int m = 0; m < k; ++m) K *= dims[axes[m]];
  if (K > (uint64_t)kMaxDiag)
    return set_error(ctx, QTNH_EUNSUPPORTED, "diag: %llu diagonal entries exceed %d", (unsigned long long)K, kMaxDiag);
  BitLayout L = bit_layout(rank, dims);
  if (L.binary) {
    DiagArgs args;
    memset(&args, 0, sizeof(args));
    memcpy(args.d, d, sizeof(c128) * K);
    int kb = 0;
    for (int m = 0; m < k; ++m)
      for (int t = L.width[axes[m]] - 1; t >= 0; --t) args.bitpos[kb++] = L.shift[axes[m]] + t;
    args.kb = kb;
    // One element per thread: the workgroups in flight then cover one
    // contiguous stretch of the state (81% of HBM peak at 2^28 and 2^30
    // elements, against 68-76% with four elements per thread 2^26+ apart).
    unsigned blocks = grid_for(n, kGateThreads, kMaxStreamBlocks);
    if (K > 4)
      hipLaunchKernelGGL(diag_bits_lds_kernel, dim3(blocks), dim3(kGateThreads), 0, ctx->stream, state, args, n);
    else
      hipLaunchKernelGGL(diag_bits_kernel, dim3(blocks), dim3(kGateThreads), 0, ctx->stream, state, args, n);
    return check_hip(ctx, hipGetLastError(), "diag launch");
  }
  DiagGenArgs args;
  memset(&args, 0, sizeof(args));
  memcpy(args.d, d, sizeof(c128) * K);
  uint64_t st[kMaxRank];
  uint64_t s = 1;
  for (int i = rank - 1; i >= 0; --i) {
    st[i] = s;
    s *= dims[i];
  }
  for (int m = 0; m < k; ++m) {
    args.stride[m] = st[axes[m]];
    args.dim[m] = dims[axes[m]];
  }
  args.k = k;
  unsigned blocks = grid_for(n, 256, 65536);
  hipLaunchKernelGGL(diag_general_kernel, dim3(blocks), dim3(256), 0, ctx->stream, state, args, n);
  return check_hip(ctx, hipGetLastError(), "diag_general launch");
}

}  // namespace qtnh_hip
