// Gate sequences on resident state vectors: qubit circuits as chains of the
// in-place gate and diagonal contractions of the contraction path.
//
// The reference applies a circuit as one Tensor::contract per gate against
// the state (examples/qft.cpp; src/tensor/contract.cpp:250-289 places each
// gate's outputs on its inputs' axes).  Here a circuit is a list of ops, each
// a dense k-qubit gate G[in][out] or a diagonal d[x], with qubit q = axis q of
// a rank-n state with every dim 2 (axis 0 is the most significant index bit).
//
// Two executors:
//  * per gate: one launch of the streaming gate / diagonal kernels per op;
//  * fused: consecutive ops are grouped into runs touching at most `tb`
//    distinct index bits.  A run is one pass over the state: each workgroup
//    gathers a tile of 2^tb amplitudes (the run's bits vary, every other bit
//    is fixed) into LDS, applies the run's ops in order, and scatters the tile
//    back.  Every op does the same arithmetic on the same values as the per
//    gate kernel (sum over inputs in row-major order from zero, products
//    without FMA), so both executors agree bit for bit; the fused one reads
//    and writes the state once per run instead of once per op.
// Batches run the k-th run of many circuits (on many states) in one launch.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "internal.hpp"

namespace qtnh_hip {

qtnh_status launch_gate(qtnh_ctx_t ctx, c128 *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                        const double *G);
qtnh_status launch_diag(qtnh_ctx_t ctx, c128 *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                        const double *d);

namespace {

typedef double double2_t __attribute__((ext_vector_type(2)));

constexpr int kMaxTileBits = 13;  // 2^13 amplitudes = 128 KiB of LDS
constexpr int kMaxOpQubits = 3;

struct TileOp {
  int kind;                 // 0: dense gate, 1: diagonal
  int kb;                   // qubits of the op inside the tile
  int pos[kMaxOpQubits];    // tile bit of op-index bit j (j = 0 most significant)
  int zpos[kMaxOpQubits];   // the same tile bits, ascending
  uint32_t coef;            // first coefficient (complex units)
  uint32_t rcode;           // register form (circuit_rb_kernel): kind | kb << 1 | reg bit j << (3 + 2 j)
  // Diagonals only: qubits outside the tile.  Byte 0: their count; byte 1 + j:
  // bit of the run's tile number holding the j-th of them (op order).  Their
  // values are fixed within a tile and pick one block of 2^kb coefficients.
  uint32_t outer;
};

// First coefficient of a diagonal's block for tile number t of its run.
__device__ __forceinline__ uint32_t outer_sel(const TileOp &op, uint64_t t) {
  const uint32_t no = op.outer & 0xffu;
  uint32_t s = 0;
  for (uint32_t j = 0; j < no; ++j) s = (s << 1) | (uint32_t)((t >> ((op.outer >> (8 + 8 * j)) & 0xffu)) & 1u);
  return s << op.kb;
}

// Coefficients of an op: 2^(2 kb) for a gate, 2^(kb + outer qubits) for a diagonal.
__device__ __forceinline__ uint32_t coef_count(const TileOp &op) {
  return op.kind == 1 ? (1u << (op.kb + (op.outer & 0xffu))) : (1u << (2 * op.kb));
}

// A phase of a run: consecutive 1- and 2-qubit ops whose bits lie in 4 tile
// bits (pmask), applied to 16 amplitudes held in registers by each thread; or
// (lds = 1) a single 3-qubit op applied in LDS.
struct PhaseDesc {
  uint32_t op_begin, op_end;
  uint32_t pmask;
  uint32_t lds;
};

struct RunDesc {
  const TileOp *ops;
  const double2_t *coef;
  double2_t *psi;
  uint64_t first;       // first tile of this run in the launch's tile numbering
  uint64_t outer_mask;  // state bits outside the tile
  uint64_t tile_mask;   // state bits of the tile
  uint32_t op_begin, op_end;
  uint32_t coef_begin;  // first coefficient of op_begin (ops' coefficients are contiguous)
  int tb;               // tile bits
  int n_outer;          // popcount(outer_mask)
  const PhaseDesc *phases;  // register phases of this run (12-bit tiles), or null
  uint32_t ph_begin, ph_end;
};

constexpr int kChunkOps = 32;     // ops staged in LDS at a time
constexpr int kChunkCoef = 512;   // their coefficients (complex)

// Parallel bit deposit: the low bits of v, in order, onto the set bits of mask.
__device__ __forceinline__ uint64_t deposit(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  while (mask) {
    const uint64_t low = mask & (0 - mask);
    if (v & 1ull) r |= low;
    v >>= 1;
    mask ^= low;
  }
  return r;
}

__device__ __forceinline__ double2_t cmul2(double2_t a, double2_t b) {
  double r, i;
  cmul_parts(a.x, a.y, b.x, b.y, &r, &i);
  return double2_t{r, i};
}

// Group gi of a k-qubit op: tile index with zero bits inserted at the op's bits.
template <int KB>
__device__ __forceinline__ int group_base(int gi, const TileOp &op) {
  int base = gi;
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    const int p = op.zpos[j];
    base = ((base >> p) << (p + 1)) | (base & ((1 << p) - 1));
  }
  return base;
}

// Dense k-qubit op on a tile in LDS.  U groups per iteration: all their loads
// are issued before any store (the groups are disjoint), so LDS latency
// overlaps across groups.
template <int KB, int NT>
__device__ void gate_in_lds(double2_t *buf, int tb, const TileOp &op, const double2_t *g) {
  constexpr int K = 1 << KB;
  constexpr int U = KB == 3 ? 1 : 2;
  int comb[K];
#pragma unroll
  for (int c = 0; c < K; ++c) {
    int off = 0;
#pragma unroll
    for (int j = 0; j < KB; ++j)
      if ((c >> (KB - 1 - j)) & 1) off |= 1 << op.pos[j];
    comb[c] = off;
  }
  const int groups = 1 << (tb - KB);
  for (int g0 = threadIdx.x; g0 < groups; g0 += NT * U) {
    // Large gates (and 2-qubit gates in 1024-thread blocks, which have 128
    // registers): re-read the coefficients from LDS in every iteration rather
    // than holding them in registers (a compiler barrier, no code).
    if (KB == 3 || (KB == 2 && NT == 1024)) asm volatile("" ::: "memory");
    int base[U];
    bool live[U];
    double2_t in[U][K];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int gi = g0 + u * NT;
      live[u] = gi < groups;
      base[u] = group_base<KB>(live[u] ? gi : g0, op);
#pragma unroll
      for (int c = 0; c < K; ++c) in[u][c] = buf[base[u] + comb[c]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!live[u]) continue;
#pragma unroll
      for (int o = 0; o < K; ++o) {
        double2_t acc = {0.0, 0.0};
#pragma unroll
        for (int i = 0; i < K; ++i) {
          const double2_t p = cmul2(in[u][i], g[i * K + o]);
          acc.x = acc.x + p.x;
          acc.y = acc.y + p.y;
        }
        buf[base[u] + comb[o]] = acc;
      }
    }
  }
}

template <int KB, int NT>
__device__ void diag_in_lds(double2_t *buf, int tb, const TileOp &op, const double2_t *d) {
  constexpr int U = 4;
  const int T = 1 << tb;
  for (int l0 = threadIdx.x; l0 < T; l0 += NT * U) {
    double2_t v[U];
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = l0 + u * NT;
      c[u] = 0;
#pragma unroll
      for (int j = 0; j < KB; ++j) c[u] = (c[u] << 1) | ((l >> op.pos[j]) & 1);
      v[u] = l < T ? buf[l] : double2_t{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int l = l0 + u * NT;
      if (l < T) buf[l] = cmul2(v[u], d[c[u]]);
    }
  }
}

// One workgroup per tile (persistent over tiles); runs[] sorted by `first`.
// Ops and their coefficients are staged into LDS in chunks; each op is
// followed by one barrier.
template <int TBMAX, int NT>
__global__ void __launch_bounds__(NT, (NT == 256 && TBMAX < 13) ? 2 : 1)
    circuit_tile_kernel(const RunDesc *__restrict__ runs, int nruns, uint64_t total_tiles) {
  constexpr int LB = NT == 1024 ? 10 : 8;  // tile bits taken from the thread index
  __shared__ double2_t buf[1 << TBMAX];
  __shared__ double2_t cbuf[kChunkCoef];
  __shared__ TileOp obuf[kChunkOps];
  __shared__ uint64_t hi_off[(1 << TBMAX) / NT + 1];
  for (uint64_t gt = blockIdx.x; gt < total_tiles; gt += gridDim.x) {
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (runs[mid].first <= gt)
        lo = mid;
      else
        hi = mid - 1;
    }
    const RunDesc &R = runs[lo];
    const uint64_t t = gt - R.first;
    const int T = 1 << R.tb;
    const int nj = (T + NT - 1) / NT;
    // Tile element l = thread + NT j: its low LB tile bits are the thread's,
    // the others come from j (table hi_off).
    uint64_t lo_mask = 0, rest = R.tile_mask;
    for (int b = 0; b < LB && rest; ++b) {
      const uint64_t low = rest & (0 - rest);
      lo_mask |= low;
      rest ^= low;
    }
    if ((int)threadIdx.x < nj) hi_off[threadIdx.x] = deposit(threadIdx.x, rest);
    double2_t *psi = R.psi + deposit(t, R.outer_mask);
    const uint64_t lo_off = deposit(threadIdx.x, lo_mask);
    __syncthreads();
    for (int l = threadIdx.x, j = 0; l < T; l += NT, ++j) buf[l] = psi[lo_off | hi_off[j]];
    uint32_t o = R.op_begin;
    uint32_t cbase = R.coef_begin;
    while (o < R.op_end) {
      // Stage the next ops whose coefficients fit the chunk.
      __syncthreads();  // previous chunk's ops are done with obuf / cbuf
      uint32_t oe = o, ce = cbase;
      while (oe < R.op_end && oe - o < (uint32_t)kChunkOps) {
        const TileOp &x = R.ops[oe];
        const uint32_t nco = coef_count(x);
        if (x.coef + nco - cbase > (uint32_t)kChunkCoef) break;
        ce = x.coef + nco;
        ++oe;
      }
      for (uint32_t i = threadIdx.x; i < oe - o; i += NT) obuf[i] = R.ops[o + i];
      for (uint32_t i = threadIdx.x; i < ce - cbase; i += NT) cbuf[i] = R.coef[cbase + i];
      __syncthreads();
      for (uint32_t i = 0; i < oe - o; ++i) {
        const TileOp &op = obuf[i];
        const double2_t *g = cbuf + (op.coef - cbase);
        if (op.kind == 1) {
          g += outer_sel(op, t);
          if (op.kb == 0)
            diag_in_lds<0, NT>(buf, R.tb, op, g);
          else if (op.kb == 1)
            diag_in_lds<1, NT>(buf, R.tb, op, g);
          else if (op.kb == 2)
            diag_in_lds<2, NT>(buf, R.tb, op, g);
          else
            diag_in_lds<3, NT>(buf, R.tb, op, g);
        } else if (op.kb == 1) {
          gate_in_lds<1, NT>(buf, R.tb, op, g);
        } else if (op.kb == 2) {
          gate_in_lds<2, NT>(buf, R.tb, op, g);
        } else {
          gate_in_lds<3, NT>(buf, R.tb, op, g);
        }
        __syncthreads();
      }
      o = oe;
      cbase = ce;
    }
    if (R.op_begin == R.op_end) __syncthreads();
    for (int l = threadIdx.x, j = 0; l < T; l += NT, ++j) psi[lo_off | hi_off[j]] = buf[l];
    __syncthreads();
  }
}

// Large states: runs whose tiles have exactly kPipeTB bits.  One 512-thread
// workgroup per CU walks its tiles with two LDS tile buffers: while the ops of
// tile k run on one buffer, tile k+1 streams into the other straight from HBM
// (global_load_lds_dwordx4: each wave instruction gathers 64 tile elements,
// 1 KiB, into consecutive LDS slots), and tile k is written back with
// non-temporal stores.  Same ops, same order, same arithmetic as
// circuit_tile_kernel.
constexpr int kPipeTB = 12, kPipeT = 1 << kPipeTB, kPipeNT = 512;
constexpr int kPipeWaves = kPipeNT / 64, kPipeCPW = kPipeT / 64 / kPipeWaves;  // 8 chunks of 64 per wave

__device__ __forceinline__ int find_run(const RunDesc *__restrict__ runs, int nruns, uint64_t gt) {
  int lo = 0, hi = nruns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (runs[mid].first <= gt)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// This lane's element offsets in tile t of run R: element l = 64 (wave + kPipeWaves k) + lane.
__device__ __forceinline__ double2_t *pipe_offsets(const RunDesc &R, uint64_t t, int wave, int lane,
                                                   uint64_t (&off)[kPipeCPW]) {
  uint64_t lo6 = 0, rest = R.tile_mask;
  for (int b = 0; b < 6; ++b) {
    const uint64_t low = rest & (0 - rest);
    lo6 |= low;
    rest ^= low;
  }
  const uint64_t lo = deposit(lane, lo6);
#pragma unroll
  for (int k = 0; k < kPipeCPW; ++k) off[k] = lo | deposit(wave + kPipeWaves * k, rest);
  return R.psi + deposit(t, R.outer_mask);
}

__device__ __forceinline__ void pipe_issue(double2_t *dstbuf, const double2_t *psi, const uint64_t (&off)[kPipeCPW],
                                           int wave) {
#pragma unroll
  for (int k = 0; k < kPipeCPW; ++k)
    __builtin_amdgcn_global_load_lds((const void *)(psi + off[k]),
                                     (void __attribute__((address_space(3))) *)(dstbuf + 64 * (wave + kPipeWaves * k)),
                                     16, 0, 0);
}

__global__ void __launch_bounds__(kPipeNT, 1)
    circuit_pipe_kernel(const RunDesc *__restrict__ runs, int nruns, uint64_t total_tiles) {
  __shared__ double2_t buf[2][kPipeT];
  __shared__ double2_t cbuf[kChunkCoef];
  __shared__ TileOp obuf[kChunkOps];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t gt = blockIdx.x;
  if (gt >= total_tiles) return;  // uniform: the whole workgroup leaves
  int r = find_run(runs, nruns, gt);
  uint64_t off[kPipeCPW];
  double2_t *psi = pipe_offsets(runs[r], gt - runs[r].first, wave, lane, off);
  pipe_issue(buf[0], psi, off, wave);
  int cur = 0;
  while (true) {
    const RunDesc &R = runs[r];
    // This tile's gather is complete (and the previous tile's stores have been
    // accepted: the counter covers both, and nothing here relies on their order).
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    // Stage the first chunk of ops before the next gather is issued (their
    // loads then do not wait behind it).
    uint32_t o = R.op_begin, cbase = R.coef_begin;
    auto stage = [&](uint32_t &oe, uint32_t &ce) {
      oe = o;
      ce = cbase;
      while (oe < R.op_end && oe - o < (uint32_t)kChunkOps) {
        const TileOp &x = R.ops[oe];
        const uint32_t nco = coef_count(x);
        if (x.coef + nco - cbase > (uint32_t)kChunkCoef) break;
        ce = x.coef + nco;
        ++oe;
      }
      for (uint32_t i = threadIdx.x; i < oe - o; i += kPipeNT) obuf[i] = R.ops[o + i];
      for (uint32_t i = threadIdx.x; i < ce - cbase; i += kPipeNT) cbuf[i] = R.coef[cbase + i];
    };
    uint32_t oe, ce;
    stage(oe, ce);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const uint64_t gn = gt + gridDim.x;
    const bool more = gn < total_tiles;  // uniform
    int rn = r;
    uint64_t offn[kPipeCPW];
    double2_t *psin = psi;
    if (more) {
      rn = find_run(runs, nruns, gn);
      psin = pipe_offsets(runs[rn], gn - runs[rn].first, wave, lane, offn);
      pipe_issue(buf[cur ^ 1], psin, offn, wave);
    }
    double2_t *tb = buf[cur];
    while (true) {
      for (uint32_t i = 0; i < oe - o; ++i) {
        const TileOp &op = obuf[i];
        const double2_t *g = cbuf + (op.coef - cbase);
        if (op.kind == 1) {
          g += outer_sel(op, gt - R.first);
          if (op.kb == 0)
            diag_in_lds<0, kPipeNT>(tb, kPipeTB, op, g);
          else if (op.kb == 1)
            diag_in_lds<1, kPipeNT>(tb, kPipeTB, op, g);
          else if (op.kb == 2)
            diag_in_lds<2, kPipeNT>(tb, kPipeTB, op, g);
          else
            diag_in_lds<3, kPipeNT>(tb, kPipeTB, op, g);
        } else if (op.kb == 1) {
          gate_in_lds<1, kPipeNT>(tb, kPipeTB, op, g);
        } else if (op.kb == 2) {
          gate_in_lds<2, kPipeNT>(tb, kPipeTB, op, g);
        } else {
          gate_in_lds<3, kPipeNT>(tb, kPipeTB, op, g);
        }
        __syncthreads();
      }
      o = oe;
      cbase = ce;
      if (o >= R.op_end) break;
      // More ops than one chunk: stage the next chunk (waits behind the gather).
      stage(oe, ce);
      __syncthreads();
    }
    // Write the tile back.
#pragma unroll
    for (int k = 0; k < kPipeCPW; ++k)
      __builtin_nontemporal_store(tb[64 * (wave + kPipeWaves * k) + lane], psi + off[k]);
    if (!more) break;
    gt = gn;
    r = rn;
    psi = psin;
#pragma unroll
    for (int k = 0; k < kPipeCPW; ++k) off[k] = offn[k];
    cur ^= 1;
  }
}

// Register-blocked runs (12-bit tiles, 256 threads): per phase, each thread
// takes the 16 tile amplitudes that differ only in the phase's 4 bits from
// LDS into registers, applies the phase's ops there (compile-time indexing:
// one instantiation per op shape and register-bit assignment) and writes them
// back; one LDS round trip and one barrier per phase instead of per op.  The
// ops do exactly the arithmetic of gate_in_lds / diag_in_lds.  LDS slots are
// XOR-swizzled (slot = l ^ ((l >> 4) & 15)) so that phases over the low tile
// bits do not fall on one bank.
// Read-only tables (ops, phases, coefficients) seen through the constant
// address space: uniform reads become scalar loads, which do not wait behind
// the vector-memory counter (and so not behind an LDS DMA in flight).
typedef const __attribute__((address_space(4))) double2_t *ccoef_t;
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *as_const(const T *p) {
  return (const __attribute__((address_space(4))) T *)(p);
}

// Field-by-field copies out of the constant address space (scalar loads).
__device__ __forceinline__ TileOp load_op(const __attribute__((address_space(4))) TileOp *p) {
  TileOp t;
  t.kind = p->kind;
  t.kb = p->kb;
#pragma unroll
  for (int j = 0; j < kMaxOpQubits; ++j) {
    t.pos[j] = p->pos[j];
    t.zpos[j] = p->zpos[j];
  }
  t.coef = p->coef;
  t.rcode = p->rcode;
  t.outer = p->outer;
  return t;
}
__device__ __forceinline__ PhaseDesc load_phase(const __attribute__((address_space(4))) PhaseDesc *p) {
  PhaseDesc d;
  d.op_begin = p->op_begin;
  d.op_end = p->op_end;
  d.pmask = p->pmask;
  d.lds = p->lds;
  return d;
}

template <int KB, int A, int B, int C>
__device__ __forceinline__ void reg_gate(double2_t (&v)[16], ccoef_t g) {
  constexpr int K = 1 << KB;
  constexpr int M = (1 << A) | (KB > 1 ? (1 << B) : 0) | (KB > 2 ? (1 << C) : 0);
#pragma unroll
  for (int r0 = 0; r0 < 16; ++r0) {
    if (r0 & M) continue;
    // 3-qubit gates: re-read the 64 coefficients per group instead of holding
    // them in registers (a compiler barrier, no code).
    if (KB == 3) asm volatile("" ::: "memory");
    int idx[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
      int x = r0;
      if ((c >> (KB - 1)) & 1) x |= 1 << A;
      if (KB > 1 && ((c >> (KB - 2)) & 1)) x |= 1 << B;
      if (KB > 2 && (c & 1)) x |= 1 << C;
      idx[c] = x;
    }
    double2_t in[K];
#pragma unroll
    for (int c = 0; c < K; ++c) in[c] = v[idx[c]];
#pragma unroll
    for (int o = 0; o < K; ++o) {
      if (KB == 3) asm volatile("" ::: "memory");  // one output's 8 coefficients live at a time
      double2_t acc = {0.0, 0.0};
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const double2_t p = cmul2(in[i], g[i * K + o]);
        acc.x = acc.x + p.x;
        acc.y = acc.y + p.y;
      }
      v[idx[o]] = acc;
    }
  }
}

template <int KB, int A, int B, int C>
__device__ __forceinline__ void reg_diag(double2_t (&v)[16], ccoef_t d) {
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    int c = KB > 0 ? (r >> A) & 1 : 0;
    if (KB > 1) c = (c << 1) | ((r >> B) & 1);
    if (KB > 2) c = (c << 1) | ((r >> C) & 1);
    v[r] = cmul2(v[r], d[c]);
  }
}

__device__ __forceinline__ void apply_reg_op(double2_t (&v)[16], uint32_t code, ccoef_t g) {
  switch (code) {
    case 2: reg_gate<1, 0, 0, 0>(v, g); break;
    case 10: reg_gate<1, 1, 0, 0>(v, g); break;
    case 18: reg_gate<1, 2, 0, 0>(v, g); break;
    case 26: reg_gate<1, 3, 0, 0>(v, g); break;
    case 36: reg_gate<2, 0, 1, 0>(v, g); break;
    case 68: reg_gate<2, 0, 2, 0>(v, g); break;
    case 100: reg_gate<2, 0, 3, 0>(v, g); break;
    case 12: reg_gate<2, 1, 0, 0>(v, g); break;
    case 76: reg_gate<2, 1, 2, 0>(v, g); break;
    case 108: reg_gate<2, 1, 3, 0>(v, g); break;
    case 20: reg_gate<2, 2, 0, 0>(v, g); break;
    case 52: reg_gate<2, 2, 1, 0>(v, g); break;
    case 116: reg_gate<2, 2, 3, 0>(v, g); break;
    case 28: reg_gate<2, 3, 0, 0>(v, g); break;
    case 60: reg_gate<2, 3, 1, 0>(v, g); break;
    case 92: reg_gate<2, 3, 2, 0>(v, g); break;
    case 1: reg_diag<0, 0, 0, 0>(v, g); break;
    case 3: reg_diag<1, 0, 0, 0>(v, g); break;
    case 11: reg_diag<1, 1, 0, 0>(v, g); break;
    case 19: reg_diag<1, 2, 0, 0>(v, g); break;
    case 27: reg_diag<1, 3, 0, 0>(v, g); break;
    case 37: reg_diag<2, 0, 1, 0>(v, g); break;
    case 69: reg_diag<2, 0, 2, 0>(v, g); break;
    case 101: reg_diag<2, 0, 3, 0>(v, g); break;
    case 13: reg_diag<2, 1, 0, 0>(v, g); break;
    case 77: reg_diag<2, 1, 2, 0>(v, g); break;
    case 109: reg_diag<2, 1, 3, 0>(v, g); break;
    case 21: reg_diag<2, 2, 0, 0>(v, g); break;
    case 53: reg_diag<2, 2, 1, 0>(v, g); break;
    case 117: reg_diag<2, 2, 3, 0>(v, g); break;
    case 29: reg_diag<2, 3, 0, 0>(v, g); break;
    case 61: reg_diag<2, 3, 1, 0>(v, g); break;
    case 93: reg_diag<2, 3, 2, 0>(v, g); break;
    // The host emits only the codes above (reg_code_ok).  With no default
    // path, each leaf of the compiler's branch tree is a case, and the 16
    // amplitudes are not copied to a merge register set on the way (32 64-bit
    // moves per thread per op, measured in the gfx950 code).
    default: __builtin_unreachable();
  }
}

__device__ __forceinline__ int rb_slot(int l) { return l ^ ((l >> 4) & 15); }

// A 3-qubit op on the swizzled 12-bit tile: the arithmetic of gate_in_lds<3> / diag_in_lds<3>.
__device__ __forceinline__ void rb_lds_op3(double2_t *buf, const TileOp &op, ccoef_t g) {
  if (op.kind == 1) {
    // The 8 coefficients are read once (uniform loads) and selected per element.
    double2_t d[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) d[c] = g[c];
    for (int l = threadIdx.x; l < 4096; l += 256) {
      const int b0 = (l >> op.pos[0]) & 1, b1 = (l >> op.pos[1]) & 1, b2 = (l >> op.pos[2]) & 1;
      const double2_t lo = b1 ? (b2 ? d[3] : d[2]) : (b2 ? d[1] : d[0]);
      const double2_t hi = b1 ? (b2 ? d[7] : d[6]) : (b2 ? d[5] : d[4]);
      buf[rb_slot(l)] = cmul2(buf[rb_slot(l)], b0 ? hi : lo);
    }
    return;
  }
  int comb[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int off = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if ((c >> (2 - j)) & 1) off |= 1 << op.pos[j];
    comb[c] = off;
  }
  for (int gi = threadIdx.x; gi < 512; gi += 256) {
    asm volatile("" ::: "memory");  // coefficients re-read per group, not held in registers
    int base = gi;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int p = op.zpos[j];
      base = ((base >> p) << (p + 1)) | (base & ((1 << p) - 1));
    }
    double2_t in[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) in[c] = buf[rb_slot(base + comb[c])];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      double2_t acc = {0.0, 0.0};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const double2_t p = cmul2(in[i], g[i * 8 + o]);
        acc.x = acc.x + p.x;
        acc.y = acc.y + p.y;
      }
      buf[rb_slot(base + comb[o])] = acc;
    }
  }
}

// Split of a 12-bit tile for the register-blocked layout: its lowest 8 state
// bits (taken from the thread index) and the other 4.
__device__ __forceinline__ void rb_masks(const RunDesc &R, uint64_t &lo_mask, uint64_t &rest) {
  lo_mask = 0;
  rest = R.tile_mask;
  for (int b = 0; b < 8; ++b) {
    const uint64_t low = rest & (0 - rest);
    lo_mask |= low;
    rest ^= low;
  }
}

// The phases of run R on tile t, held in the swizzled buffer; ends with a barrier.
__device__ __forceinline__ void rb_phases(double2_t *buf, const RunDesc &R, uint64_t t, int tid) {
  const auto phases = as_const(R.phases);
  const auto ops = as_const(R.ops);
  const ccoef_t coef = as_const(R.coef);
  for (uint32_t ph = R.ph_begin; ph < R.ph_end; ++ph) {
    const PhaseDesc P = load_phase(phases + ph);
    if (P.lds) {
      // A 3-qubit op, in LDS (its tile bits index the swizzled buffer through rb_slot).
      const TileOp op = load_op(ops + P.op_begin);
      rb_lds_op3(buf, op, coef + op.coef + (op.kind == 1 ? outer_sel(op, t) : 0u));
      lds_barrier();
      continue;
    }
    // The phase's 4 tile bits as values (ascending): register index r sets bit
    // k -> tile bit pb[k].  The thread's base element is tid with zero bits
    // inserted at those 4 positions (the tile's other 8 bits, in order).
    int pb[4];
    uint32_t pm = P.pmask;
    int base = tid;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      pb[k] = (int)(pm & (0u - pm));
      pm ^= (uint32_t)pb[k];
      base = ((base & ~(pb[k] - 1)) << 1) | (base & (pb[k] - 1));
    }
    // rb_slot is linear over XOR and base has no phase bit set, so each slot is
    // rb_slot(base) XOR a uniform term.
    const int slot0 = rb_slot(base);
    int slot[16];
#pragma unroll
    for (int r = 0; r < 16; ++r)
      slot[r] = slot0 ^ rb_slot(((r & 1) ? pb[0] : 0) | ((r & 2) ? pb[1] : 0) | ((r & 4) ? pb[2] : 0) |
                                ((r & 8) ? pb[3] : 0));
    double2_t v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = buf[slot[r]];
    // The next op's descriptor is requested before the current op runs.
    TileOp op = load_op(ops + P.op_begin);
    for (uint32_t o = P.op_begin; o < P.op_end; ++o) {
      const TileOp nxt = load_op(ops + (o + 1 < P.op_end ? o + 1 : o));
      apply_reg_op(v, op.rcode, coef + op.coef + (op.kind == 1 ? outer_sel(op, t) : 0u));
      op = nxt;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) buf[slot[r]] = v[r];
    lds_barrier();
  }
}

// This thread's 16 elements of tile t of run R (element l = tid + 256 j) and
// their global offsets.
struct RbTileRefs {
  __attribute__((address_space(1))) double2_t *psi;
  uint64_t lo_off;
  uint64_t hb[4];  // state-bit values of the tile's 4 upper bits, ascending
  // Offset of element tid + 256 j.
  __device__ __forceinline__ uint64_t off(int j) const {
    return lo_off | ((j & 1) ? hb[0] : 0) | ((j & 2) ? hb[1] : 0) | ((j & 4) ? hb[2] : 0) | ((j & 8) ? hb[3] : 0);
  }
};
__device__ __forceinline__ RbTileRefs rb_refs(const RunDesc &R, uint64_t t, int tid) {
  uint64_t lo_mask, rest;
  rb_masks(R, lo_mask, rest);
  RbTileRefs f;
  f.psi = (__attribute__((address_space(1))) double2_t *)(R.psi + deposit(t, R.outer_mask));
  f.lo_off = deposit(tid, lo_mask);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f.hb[k] = rest & (0 - rest);
    rest ^= f.hb[k];
  }
  return f;
}

// Write tile t of run R back from the swizzled buffer (element l = tid + 256 j).
// Global address space: the stores cannot alias LDS.
__device__ __forceinline__ void rb_store(const double2_t *buf, const RbTileRefs &f, int tid) {
#pragma unroll
  for (int j = 0; j < 16; ++j) __builtin_nontemporal_store(buf[rb_slot(tid + 256 * j)], f.psi + f.off(j));
}

__global__ void __launch_bounds__(256, 2) circuit_rb_kernel(const RunDesc *__restrict__ runs, int nruns,
                                                            uint64_t total_tiles) {
  __shared__ double2_t buf[1 << 12];
  const int tid = threadIdx.x;
  for (uint64_t gt = blockIdx.x; gt < total_tiles; gt += gridDim.x) {
    const RunDesc &R = runs[find_run(runs, nruns, gt)];
    const uint64_t t = gt - R.first;
    // Gather: element l = tid + 256 j (its low 8 tile bits are the thread's).
    const RbTileRefs f = rb_refs(R, t, tid);
#pragma unroll
    for (int j = 0; j < 16; ++j) buf[rb_slot(tid + 256 * j)] = __builtin_nontemporal_load(f.psi + f.off(j));
    __syncthreads();
    rb_phases(buf, R, t, tid);
    rb_store(buf, f, tid);
    __syncthreads();
  }
}

// Register-blocked runs with the next tile's loads in flight: as
// circuit_rb_kernel (two 256-thread workgroups per CU, one 64 KiB tile each),
// but each thread issues the loads of its 16 elements of the next tile into
// registers before the phases of the current tile, so HBM reads overlap the
// phases (the LDS-only barriers between phases do not wait for them), and
// moves them into the buffer after the current tile's write-back.  Same
// ops, same order, same arithmetic as circuit_rb_kernel.
__global__ void __launch_bounds__(256, 2) circuit_rbp_kernel(const RunDesc *__restrict__ runs, int nruns,
                                                             uint64_t total_tiles) {
  __shared__ double2_t buf[1 << 12];
  const int tid = threadIdx.x;
  uint64_t gt = blockIdx.x;
  if (gt >= total_tiles) return;  // uniform: the whole workgroup leaves
  double2_t pre[16];
  {
    const RunDesc &R = runs[find_run(runs, nruns, gt)];
    const RbTileRefs f = rb_refs(R, gt - R.first, tid);
#pragma unroll
    for (int j = 0; j < 16; ++j) pre[j] = __builtin_nontemporal_load(f.psi + f.off(j));
#pragma unroll
    for (int j = 0; j < 16; ++j) buf[rb_slot(tid + 256 * j)] = pre[j];
    lds_barrier();
  }
  while (true) {
    const uint64_t gn = gt + gridDim.x;
    const bool more = gn < total_tiles;  // uniform
    if (more) {
      const RunDesc &Rn = runs[find_run(runs, nruns, gn)];
      const RbTileRefs f = rb_refs(Rn, gn - Rn.first, tid);
#pragma unroll
      for (int j = 0; j < 16; ++j) pre[j] = __builtin_nontemporal_load(f.psi + f.off(j));
    }
    const RunDesc &R = runs[find_run(runs, nruns, gt)];
    const uint64_t t = gt - R.first;
    rb_phases(buf, R, t, tid);
    rb_store(buf, rb_refs(R, t, tid), tid);
    if (!more) break;
    // Every wave has read the buffer for its stores: refill it with the next
    // tile (waiting here for its loads only, issued before the stores).
    lds_barrier();
#pragma unroll
    for (int j = 0; j < 16; ++j) buf[rb_slot(tid + 256 * j)] = pre[j];
    lds_barrier();
    gt = gn;
  }
}

// ---- <Z_q> for a list of qubits --------------------------------------------
// Chunks of 2^kChunkBits consecutive amplitudes.  Per chunk: the total
// probability and, for each of the chunk's own index bits, the probability
// with that bit set.  Bits above the chunk are constant inside it.
constexpr int kChunkBits = 12;
constexpr int kZThreads = 256;
constexpr int kZSlots = 1 + kChunkBits;

__global__ void __launch_bounds__(kZThreads) zchunk_kernel(const double2_t *__restrict__ psi, uint64_t n, int cb,
                                                           double *__restrict__ part, uint64_t nchunks) {
  __shared__ double red[kZThreads];
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    // Thread t owns chunk elements t + kZThreads j: bits 0..7 of the element are t's bits.
    double tot = 0.0, hib[kChunkBits];
#pragma unroll
    for (int b = 0; b < kChunkBits; ++b) hib[b] = 0.0;
    const uint64_t e0 = ch << cb;
    const int csize = 1 << cb;
    for (int l = threadIdx.x; l < csize; l += kZThreads) {
      const uint64_t e = e0 + l;
      double p = 0.0;
      if (e < n) {
        const double2_t v = psi[e];
        p = v.x * v.x;
        p = p + v.y * v.y;
      }
      tot = tot + p;
#pragma unroll
      for (int b = 8; b < kChunkBits; ++b)
        if ((l >> b) & 1) hib[b] = hib[b] + p;
    }
    // slot 0: total; slot 1 + b: probability with bit b set.
#pragma unroll
    for (int s = 0; s < kZSlots; ++s) {
      double v;
      if (s == 0)
        v = tot;
      else if (s - 1 < 8)
        v = ((threadIdx.x >> (s - 1)) & 1) ? tot : 0.0;
      else
        v = hib[s - 1];
      red[threadIdx.x] = v;
      __syncthreads();
      for (int w = kZThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) part[ch * kZSlots + s] = red[0];
      __syncthreads();
    }
  }
}

// One workgroup per requested qubit: <Z_b> = sum over chunks of (total - 2 set_b).
__global__ void __launch_bounds__(kZThreads) zfinal_kernel(const double *__restrict__ part, uint64_t nchunks, int cb,
                                                           const int *__restrict__ bits, double *__restrict__ out) {
  __shared__ double red[kZThreads];
  const int b = bits[blockIdx.x];
  double s = 0.0;
  for (uint64_t ch = threadIdx.x; ch < nchunks; ch += kZThreads) {
    const double tot = part[ch * kZSlots];
    double v;
    if (b < cb)
      v = tot - 2.0 * part[ch * kZSlots + 1 + b];
    else
      v = (((ch << cb) >> b) & 1) ? -tot : tot;
    s = s + v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kZThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace
}  // namespace qtnh_hip

using qtnh_hip::c128;
using qtnh_hip::RunDesc;
using qtnh_hip::TileOp;

// ---- host side ---------------------------------------------------------------
struct qtnh_circuit_s {
  struct Op {
    int kind;  // 0 gate, 1 diagonal
    int k;
    uint16_t q[3];
    std::vector<double> coef;  // interleaved complex: gates (2^k)^2, diagonals 2^k
  };
  int n = 0;
  std::vector<Op> ops;
  // Tile shape chosen for this circuit by qtnh_circuit_tune (0: the context's).
  int tuned_tb = 0, tuned_low = 0;
  // Compiled form (device), rebuilt when ops or the tile size change.
  int compiled_tb = -1;
  int device = -1;
  size_t compiled_ops = 0;
  TileOp *d_ops = nullptr;
  qtnh_hip::double2_t *d_coef = nullptr;
  struct Run {
    uint32_t begin, end;
    std::vector<int> tbits;  // ascending state bits
  };
  std::vector<Run> runs;
  std::vector<TileOp> tops;  // host copy of the compiled ops
  qtnh_hip::PhaseDesc *d_phases = nullptr;
  std::vector<std::pair<uint32_t, uint32_t>> run_phases;  // per run: [begin, end) into the phases
  // Uploaded run descriptors of fused runs, per state buffer: a repeated run
  // on the same state launches kernels only (no copy), so it can be captured
  // into a graph.  Cleared when the circuit is recompiled.
  struct DescCache {
    RunDesc *d = nullptr;
    std::vector<RunDesc> h;
    std::vector<size_t> starts;
  };
  std::map<const void *, DescCache> desc_cache;
  // Device memory of earlier compiled forms of the same ops (another tile
  // shape or device).  A graph captured from an earlier fused run may still
  // read it, so it is freed only when ops are appended or the circuit is
  // destroyed.
  std::vector<void *> retired;
};

namespace qtnh_hip {
namespace {

int g_tile_bits = 12;
// Large states with 12-bit tiles: 2 register-blocked kernel, 1 pipelined LDS
// kernel, 0 plain LDS kernel.
int g_circuit_mode = 2;
int g_low_bits = 4;  // lowest state bits in every tile: 2^4 x 16 B = 256 B contiguous rows

// State bits an op needs inside the tile: all of a gate's; none of a
// diagonal's (a qubit outside the tile is fixed per tile and selects the
// diagonal's coefficients there).
uint64_t mask_of(const qtnh_circuit_s &c, const qtnh_circuit_s::Op &op) {
  uint64_t m = 0;
  if (op.kind == 1) return 0;
  for (int j = 0; j < op.k; ++j) m |= 1ull << (c.n - 1 - op.q[j]);
  return m;
}

// Greedy partition of the op sequence into runs of at most tb distinct gate
// bits, each widened with the lowest free bits (coalesced tile rows) up to
// tb.  Runs are contiguous in op order, so every amplitude sees the ops in
// program order; with that constraint the greedy (longest feasible run first)
// gives the fewest runs.
void plan_runs(qtnh_circuit_s &c, int tb, int low_bits) {
  c.runs.clear();
  const int tbe = std::min(tb, c.n);
  const uint64_t low = (1ull << std::min(std::min(low_bits, tbe - 3 > 0 ? tbe - 3 : 0), c.n)) - 1;
  uint64_t cur = low;
  uint32_t begin = 0;
  auto close = [&](uint32_t end) {
    if (end == begin) return;
    uint64_t m = cur;
    for (int b = 0; b < c.n && __builtin_popcountll(m) < tbe; ++b) m |= 1ull << b;
    qtnh_circuit_s::Run r;
    r.begin = begin;
    r.end = end;
    for (int b = 0; b < c.n; ++b)
      if ((m >> b) & 1) r.tbits.push_back(b);
    c.runs.push_back(r);
  };
  for (uint32_t i = 0; i < c.ops.size(); ++i) {
    const uint64_t m = mask_of(c, c.ops[i]);
    if (__builtin_popcountll(cur | m) > tbe) {
      close(i);
      begin = i;
      cur = low;
    }
    cur |= m;
  }
  close((uint32_t)c.ops.size());
}

// Detaches the current compiled form; its device memory moves to `retired`.
void retire_compiled(qtnh_circuit_s &c) {
  for (auto &kv : c.desc_cache)
    if (kv.second.d) c.retired.push_back(kv.second.d);
  c.desc_cache.clear();
  for (void *p : {(void *)c.d_ops, (void *)c.d_coef, (void *)c.d_phases})
    if (p) c.retired.push_back(p);
  c.d_ops = nullptr;
  c.d_coef = nullptr;
  c.d_phases = nullptr;
  c.compiled_tb = -1;
}

// Frees the current and every retired compiled form (the caller has drained
// the streams that used them).
void free_compiled(qtnh_circuit_s &c) {
  retire_compiled(c);
  for (void *p : c.retired) (void)hipFree(p);
  c.retired.clear();
}

// Whether a register-phase op code is one of apply_reg_op's cases: a gate on
// 1 or 2 register bits, or a diagonal on 0-2 of them, bits distinct.
bool reg_code_ok(uint32_t code) {
  const uint32_t kind = code & 1u, kb = (code >> 1) & 3u;
  if (kb > 2 || (kind == 0 && kb == 0) || (code >> (3 + 2 * kb)) != 0) return false;
  return kb < 2 || ((code >> 3) & 3u) != ((code >> 5) & 3u);
}

// Register phases of a 12-bit-tile run: consecutive ops whose tile bits fit 4,
// padded with the highest free tile bits (threads then vary in the low bits).
void plan_phases(std::vector<TileOp> &tops, uint32_t begin, uint32_t end, std::vector<PhaseDesc> &out) {
  uint32_t cur = 0, pb = begin;
  auto mask_of_op = [&](const TileOp &t) {
    uint32_t m = 0;
    for (int j = 0; j < t.kb; ++j) m |= 1u << t.pos[j];
    return m;
  };
  auto close = [&](uint32_t pe) {
    if (pe == pb) return;
    if (tops[pb].kb == 3) {
      PhaseDesc ph;
      ph.op_begin = pb;
      ph.op_end = pe;
      ph.pmask = 0;
      ph.lds = 1;
      tops[pb].rcode = 0;
      out.push_back(ph);
      return;
    }
    uint32_t m = cur;
    for (int b = 11; b >= 0 && __builtin_popcount(m) < 4; --b) m |= 1u << b;
    PhaseDesc ph;
    ph.op_begin = pb;
    ph.op_end = pe;
    ph.pmask = m;
    ph.lds = 0;
    // Register bit of tile bit p: its rank among the phase's 4 bits.
    auto regbit = [&](int p) { return __builtin_popcount(m & ((1u << p) - 1)); };
    for (uint32_t i = pb; i < pe; ++i) {
      TileOp &t = tops[i];
      uint32_t code = (uint32_t)t.kind | ((uint32_t)t.kb << 1);
      for (int j = 0; j < t.kb; ++j) code |= (uint32_t)regbit(t.pos[j]) << (3 + 2 * j);
      t.rcode = code;
    }
    out.push_back(ph);
  };
  for (uint32_t i = begin; i < end; ++i) {
    const uint32_t m = mask_of_op(tops[i]);
    if (tops[i].kb == 3) {
      close(i);
      pb = i;
      close(i + 1);
      pb = i + 1;
      cur = 0;
      continue;
    }
    if (__builtin_popcount(cur | m) > 4) {
      close(i);
      pb = i;
      cur = 0;
    }
    cur |= m;
  }
  close(end);
}

// Tile shape for an n-qubit circuit: the autotuned one of this context, else the process default.
void tile_shape(qtnh_ctx_t ctx, int n, int *tb, int *low) {
  const bool tuned = n < 64 && ctx->circ_tb[n] > 0;
  *tb = tuned ? ctx->circ_tb[n] : g_tile_bits;
  *low = tuned ? ctx->circ_low[n] : g_low_bits;
}

qtnh_status compile(qtnh_ctx_t ctx, qtnh_circuit_s &c) {
  int tb, low;
  tile_shape(ctx, c.n, &tb, &low);
  if (c.tuned_tb > 0) {
    tb = c.tuned_tb;
    low = c.tuned_low;
  }
  if (c.compiled_tb == tb * 16 + low && c.device == ctx->device && c.compiled_ops == c.ops.size()) return QTNH_OK;
  if (c.d_ops || c.d_coef) {
    if (c.compiled_ops == c.ops.size()) {
      // Same ops, another tile shape or device: graphs captured from the
      // earlier form stay valid, so its memory is kept.
      retire_compiled(c);
    } else {
      qtnh_status st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "circuit recompile sync");
      if (st) return st;
      free_compiled(c);
    }
  }
  plan_runs(c, tb, low);
  std::vector<TileOp> tops(std::max<size_t>(c.ops.size(), 1));
  std::vector<double> coef;
  for (const auto &r : c.runs) {
    for (uint32_t i = r.begin; i < r.end; ++i) {
      const auto &op = c.ops[i];
      TileOp t;
      memset(&t, 0, sizeof(t));
      t.kind = op.kind;
      t.coef = (uint32_t)(coef.size() / 2);
      // Op-index bits j inside the tile (tj) and outside it (oj, diagonals only).
      int tj[3], oj[3], nt = 0, no = 0;
      for (int j = 0; j < op.k; ++j) {
        const int sb = c.n - 1 - op.q[j];
        const auto it = std::find(r.tbits.begin(), r.tbits.end(), sb);
        if (it != r.tbits.end()) {
          t.pos[nt] = (int)(it - r.tbits.begin());
          t.zpos[nt] = t.pos[nt];
          tj[nt++] = j;
        } else {
          // Bit of the tile number: the state bit's rank among the bits outside the tile.
          const int below = (int)(std::lower_bound(r.tbits.begin(), r.tbits.end(), sb) - r.tbits.begin());
          t.outer |= (uint32_t)(sb - below) << (8 + 8 * no);
          oj[no++] = j;
        }
      }
      if (op.kind == 0 && no) return set_error(ctx, QTNH_EINVAL, "circuit: gate outside its tile (planner)");
      t.kb = nt;
      t.outer |= (uint32_t)no;
      std::sort(t.zpos, t.zpos + nt);
      if (op.kind == 0 || no == 0) {
        coef.insert(coef.end(), op.coef.begin(), op.coef.end());
      } else {
        // Block o (outer bits, op order) holds the 2^nt coefficients of the tile bits.
        for (int o = 0; o < (1 << no); ++o)
          for (int tt = 0; tt < (1 << nt); ++tt) {
            int idx = 0;
            for (int m = 0; m < nt; ++m)
              if ((tt >> (nt - 1 - m)) & 1) idx |= 1 << (op.k - 1 - tj[m]);
            for (int m = 0; m < no; ++m)
              if ((o >> (no - 1 - m)) & 1) idx |= 1 << (op.k - 1 - oj[m]);
            coef.push_back(op.coef[2 * idx]);
            coef.push_back(op.coef[2 * idx + 1]);
          }
      }
      tops[i] = t;
    }
  }
  if (coef.empty()) coef.assign(2, 0.0);
  std::vector<PhaseDesc> phases;
  c.run_phases.clear();
  for (const auto &r : c.runs) {
    const uint32_t p0 = (uint32_t)phases.size();
    if (r.tbits.size() == 12) plan_phases(tops, r.begin, r.end, phases);
    c.run_phases.push_back({p0, (uint32_t)phases.size()});
  }
  for (const auto &ph : phases)
    for (uint32_t i = ph.op_begin; !ph.lds && i < ph.op_end; ++i)
      if (!reg_code_ok(tops[i].rcode)) return set_error(ctx, QTNH_EINVAL, "circuit: register op form (planner)");
  if (phases.empty()) phases.resize(1);
  qtnh_status st = check_hip(ctx, hipMalloc((void **)&c.d_ops, tops.size() * sizeof(TileOp)), "circuit alloc");
  if (!st) st = check_hip(ctx, hipMalloc((void **)&c.d_phases, phases.size() * sizeof(PhaseDesc)), "circuit alloc");
  if (!st)
    st = check_hip(ctx, hipMemcpyAsync(c.d_phases, phases.data(), phases.size() * sizeof(PhaseDesc),
                                       hipMemcpyHostToDevice, ctx->stream),
                   "circuit upload");
  if (!st) st = check_hip(ctx, hipMalloc((void **)&c.d_coef, coef.size() * sizeof(double)), "circuit alloc");
  if (!st)
    st = check_hip(ctx, hipMemcpyAsync(c.d_ops, tops.data(), tops.size() * sizeof(TileOp), hipMemcpyHostToDevice,
                                       ctx->stream),
                   "circuit upload");
  if (!st)
    st = check_hip(ctx, hipMemcpyAsync(c.d_coef, coef.data(), coef.size() * sizeof(double), hipMemcpyHostToDevice,
                                       ctx->stream),
                   "circuit upload");
  if (st) {
    // Only this attempt's buffers: no launch has used them (retired forms stay).
    (void)hipStreamSynchronize(ctx->stream);
    for (void *p : {(void *)c.d_ops, (void *)c.d_coef, (void *)c.d_phases})
      if (p) (void)hipFree(p);
    c.d_ops = nullptr;
    c.d_coef = nullptr;
    c.d_phases = nullptr;
    c.compiled_tb = -1;
    return st;
  }
  c.tops = tops;
  c.compiled_tb = tb * 16 + low;
  c.device = ctx->device;
  c.compiled_ops = c.ops.size();
  return QTNH_OK;
}

RunDesc make_desc(const qtnh_circuit_s &c, size_t ri, c128 *psi, uint64_t first) {
  const qtnh_circuit_s::Run &r = c.runs[ri];
  RunDesc d;
  memset(&d, 0, sizeof(d));
  d.phases = c.d_phases;
  d.ph_begin = c.run_phases[ri].first;
  d.ph_end = c.run_phases[ri].second;
  d.ops = c.d_ops;
  d.coef = c.d_coef;
  d.psi = (double2_t *)psi;
  d.first = first;
  d.op_begin = r.begin;
  d.op_end = r.end;
  d.coef_begin = r.begin < c.tops.size() ? c.tops[r.begin].coef : 0;
  d.tb = (int)r.tbits.size();
  uint64_t tile = 0;
  for (size_t j = 0; j < r.tbits.size(); ++j) tile |= 1ull << r.tbits[j];
  d.tile_mask = tile;
  const uint64_t all = c.n == 64 ? ~0ull : ((1ull << c.n) - 1);
  d.outer_mask = all & ~tile;
  d.n_outer = __builtin_popcountll(d.outer_mask);
  return d;
}

uint64_t desc_tiles(const RunDesc &d) { return 1ull << d.n_outer; }

// Launches groups of runs: group g covers descs [starts[g], starts[g+1]), one
// launch each, in order.  The descriptors go to the device in one copy.
// Launches the run groups [starts[g], starts[g+1]) of descriptors already on
// the device (dd); descs is their host copy.
qtnh_status launch_groups(qtnh_ctx_t ctx, const RunDesc *dd, const std::vector<RunDesc> &descs,
                          const std::vector<size_t> &starts) {
  qtnh_status st = QTNH_OK;
  for (size_t g = 0; g + 1 < starts.size(); ++g) {
    const size_t b = starts[g], e = starts[g + 1];
    if (b == e) continue;
    int tbmax = 0;
    uint64_t total = 0;
    for (size_t i = b; i < e; ++i) {
      tbmax = std::max(tbmax, descs[i].tb);
      total = descs[i].first + desc_tiles(descs[i]);
    }
    const int nr = (int)(e - b);
    bool all12 = true;
    for (size_t i = b; i < e; ++i) all12 = all12 && descs[i].tb == kPipeTB && descs[i].ph_end > descs[i].ph_begin;
    if (all12 && total >= 2ull * ctx->num_cus && g_circuit_mode == 3) {
      const unsigned rbk = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus * 2);
      hipLaunchKernelGGL(circuit_rbp_kernel, dim3(rbk), dim3(256), 0, ctx->stream, dd + b, nr, total);
      st = check_hip(ctx, hipGetLastError(), "circuit pipelined register-blocked launch");
      if (st) return st;
      continue;
    }
    // Register-blocked for any number of tiles: one barrier per phase instead of
    // one per op also shortens single-tile runs (QFT-12: 90 us instead of 115).
    if (all12 && g_circuit_mode == 2) {
      const unsigned rbk = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus * 2);
      hipLaunchKernelGGL(circuit_rb_kernel, dim3(rbk), dim3(256), 0, ctx->stream, dd + b, nr, total);
      st = check_hip(ctx, hipGetLastError(), "circuit register-blocked launch");
      if (st) return st;
      continue;
    }
    if (all12 && total >= 4ull * ctx->num_cus && g_circuit_mode == 1) {
      const unsigned pb = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus);
      hipLaunchKernelGGL(circuit_pipe_kernel, dim3(pb), dim3(kPipeNT), 0, ctx->stream, dd + b, nr, total);
      st = check_hip(ctx, hipGetLastError(), "circuit pipe launch");
      if (st) return st;
      continue;
    }
    // Few tiles: 1024 threads per tile; many: 256 threads, several tiles per CU.
    const bool wide = total < (uint64_t)ctx->num_cus;
    const unsigned per_cu = wide ? 1 : tbmax <= 10 ? 8 : tbmax == 11 ? 4 : tbmax == 12 ? 2 : 1;
    const unsigned blocks = (unsigned)std::min<uint64_t>(total, (uint64_t)ctx->num_cus * per_cu);
#define QTNH_TILE_LAUNCH(TBM)                                                                                       \
  if (wide)                                                                                                       \
    hipLaunchKernelGGL((circuit_tile_kernel<TBM, 1024>), dim3(blocks), dim3(1024), 0, ctx->stream, dd + b, nr, total); \
  else                                                                                                            \
    hipLaunchKernelGGL((circuit_tile_kernel<TBM, 256>), dim3(blocks), dim3(256), 0, ctx->stream, dd + b, nr, total);
    if (tbmax <= 10) {
      QTNH_TILE_LAUNCH(10)
    } else if (tbmax == 11) {
      QTNH_TILE_LAUNCH(11)
    } else if (tbmax == 12) {
      QTNH_TILE_LAUNCH(12)
    } else {
      QTNH_TILE_LAUNCH(13)
    }
#undef QTNH_TILE_LAUNCH
    st = check_hip(ctx, hipGetLastError(), "circuit tile launch");
    if (st) return st;
  }
  return QTNH_OK;
}

// Uploads the descriptors into the context's descriptor buffer, then launches.
qtnh_status launch_run_groups(qtnh_ctx_t ctx, const std::vector<RunDesc> &descs, const std::vector<size_t> &starts) {
  if (descs.empty()) return QTNH_OK;
  const uint64_t bytes = descs.size() * sizeof(RunDesc);
  qtnh_status st = grow_buffer(ctx, &ctx->desc, &ctx->desc_bytes, bytes);
  if (st) return st;
  st = check_hip(ctx, hipMemcpyAsync(ctx->desc, descs.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                 "run descriptor upload");
  if (st) return st;
  return launch_groups(ctx, (const RunDesc *)ctx->desc, descs, starts);
}

constexpr size_t kDescCacheMax = 64;  // states per circuit with cached descriptors

qtnh_status run_fused(qtnh_ctx_t ctx, qtnh_circuit_s &c, c128 *psi) {
  qtnh_status st = compile(ctx, c);
  if (st) return st;
  auto hit = c.desc_cache.find(psi);
  if (hit != c.desc_cache.end()) return launch_groups(ctx, hit->second.d, hit->second.h, hit->second.starts);
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cap) != hipSuccess) {
    (void)hipGetLastError();
    cap = hipStreamCaptureStatusNone;
  }
  if (cap != hipStreamCaptureStatusNone)
    return set_error(ctx, QTNH_EUNSUPPORTED,
                     "circuit_run: first fused run of this circuit on this state during graph capture "
                     "(run it once on the state before capturing)");
  qtnh_circuit_s::DescCache dc;
  for (size_t ri = 0; ri < c.runs.size(); ++ri) {
    dc.starts.push_back(dc.h.size());
    dc.h.push_back(make_desc(c, ri, psi, 0));
  }
  dc.starts.push_back(dc.h.size());
  if (dc.h.empty()) return QTNH_OK;
  if (c.desc_cache.size() >= kDescCacheMax) return launch_run_groups(ctx, dc.h, dc.starts);
  const uint64_t bytes = dc.h.size() * sizeof(RunDesc);
  st = check_hip(ctx, hipMalloc((void **)&dc.d, bytes), "run descriptor alloc");
  if (st) return st;
  auto &e = c.desc_cache[psi];
  e = std::move(dc);
  st = check_hip(ctx, hipMemcpyAsync(e.d, e.h.data(), bytes, hipMemcpyHostToDevice, ctx->stream),
                 "run descriptor upload");
  if (st) {
    (void)hipFree(e.d);
    c.desc_cache.erase(psi);
    return st;
  }
  return launch_groups(ctx, e.d, e.h, e.starts);
}

qtnh_status run_per_gate(qtnh_ctx_t ctx, const qtnh_circuit_s &c, c128 *psi) {
  std::vector<uint64_t> dims(c.n, 2);
  for (const auto &op : c.ops) {
    qtnh_status st = op.kind == 0 ? launch_gate(ctx, psi, c.n, dims.data(), op.k, op.q, op.coef.data())
                                  : launch_diag(ctx, psi, c.n, dims.data(), op.k, op.q, op.coef.data());
    if (st) return st;
  }
  return QTNH_OK;
}

// ---- seeded generators ----------------------------------------------------
uint64_t mix64(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct HostRng {
  uint64_t seed, k = 0;
  explicit HostRng(uint64_t s) : seed(s) {}
  uint64_t next() { return mix64(seed, k++); }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  void normal2(double *a, double *b) {
    const double u1 = (double)((next() >> 11) + 1) * 0x1.0p-53, u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    *a = r * std::cos(2.0 * M_PI * u2);
    *b = r * std::sin(2.0 * M_PI * u2);
  }
};

// Haar-random SU(2) (Mezzadri 2007): QR of a complex Gaussian 2x2 with the
// phases of R's diagonal moved into Q, then divided by sqrt(det).
// Returns G[in][out] = U[out][in].
void random_su2(HostRng &rng, double *G) {
  double z[4][2];
  for (int i = 0; i < 4; ++i) rng.normal2(&z[i][0], &z[i][1]);
  // columns a = (z00, z10), b = (z01, z11) of Z
  const double ar[2] = {z[0][0], z[2][0]}, ai[2] = {z[0][1], z[2][1]};
  const double br[2] = {z[1][0], z[3][0]}, bi[2] = {z[1][1], z[3][1]};
  const double na = std::sqrt(ar[0] * ar[0] + ai[0] * ai[0] + ar[1] * ar[1] + ai[1] * ai[1]);
  double q0r[2], q0i[2];
  for (int i = 0; i < 2; ++i) {
    q0r[i] = ar[i] / na;
    q0i[i] = ai[i] / na;
  }
  // r01 = <q0, b>; b' = b - q0 r01
  double r01r = 0, r01i = 0;
  for (int i = 0; i < 2; ++i) {
    r01r += q0r[i] * br[i] + q0i[i] * bi[i];
    r01i += q0r[i] * bi[i] - q0i[i] * br[i];
  }
  double cr[2], ci[2];
  for (int i = 0; i < 2; ++i) {
    cr[i] = br[i] - (q0r[i] * r01r - q0i[i] * r01i);
    ci[i] = bi[i] - (q0r[i] * r01i + q0i[i] * r01r);
  }
  const double nc = std::sqrt(cr[0] * cr[0] + ci[0] * ci[0] + cr[1] * cr[1] + ci[1] * ci[1]);
  double q1r[2], q1i[2];
  for (int i = 0; i < 2; ++i) {
    q1r[i] = cr[i] / nc;
    q1i[i] = ci[i] / nc;
  }
  // R's diagonal (na, nc) is real positive, so Q is already phase-fixed.
  // U = Q / sqrt(det Q)
  const double dr = (q0r[0] * q1r[1] - q0i[0] * q1i[1]) - (q1r[0] * q0r[1] - q1i[0] * q0i[1]);
  const double di = (q0r[0] * q1i[1] + q0i[0] * q1r[1]) - (q1r[0] * q0i[1] + q1i[0] * q0r[1]);
  const double ph = std::atan2(di, dr) / 2.0;
  const double sr = std::cos(-ph), si = std::sin(-ph);
  double U[2][2][2];  // U[row][col][re/im]
  for (int i = 0; i < 2; ++i) {
    U[i][0][0] = q0r[i] * sr - q0i[i] * si;
    U[i][0][1] = q0r[i] * si + q0i[i] * sr;
    U[i][1][0] = q1r[i] * sr - q1i[i] * si;
    U[i][1][1] = q1r[i] * si + q1i[i] * sr;
  }
  for (int in = 0; in < 2; ++in)
    for (int out = 0; out < 2; ++out) {
      G[2 * (in * 2 + out)] = U[out][in][0];
      G[2 * (in * 2 + out) + 1] = U[out][in][1];
    }
}

bool qubits_valid(int n, int k, const uint16_t *q) {
  if (k < 1 || k > kMaxOpQubits) return false;
  for (int j = 0; j < k; ++j) {
    if (q[j] >= n) return false;
    for (int i = 0; i < j; ++i)
      if (q[i] == q[j]) return false;
  }
  return true;
}

}  // namespace
}  // namespace qtnh_hip

namespace qtnh_hip {

qtnh_status circuit_run_count(qtnh_ctx_t ctx, qtnh_circuit_t c, int *nruns) {
  qtnh_status st = compile(ctx, *c);
  if (st) return st;
  *nruns = (int)c->runs.size();
  return QTNH_OK;
}

qtnh_status circuit_advance(qtnh_ctx_t ctx, int count, const qtnh_circuit_t *cs, void *const *states,
                            const int *first_run, const int *nrun) {
  int kmax = 0;
  for (int i = 0; i < count; ++i) {
    qtnh_status st = compile(ctx, *cs[i]);
    if (st) return st;
    if (first_run[i] < 0 || nrun[i] < 0 || (size_t)first_run[i] + (size_t)nrun[i] > cs[i]->runs.size())
      return set_error(ctx, QTNH_EINVAL, "circuit advance: runs out of range");
    kmax = std::max(kmax, nrun[i]);
  }
  std::vector<RunDesc> descs;
  std::vector<size_t> starts;
  for (int k = 0; k < kmax; ++k) {
    starts.push_back(descs.size());
    uint64_t first = 0;
    for (int i = 0; i < count; ++i) {
      if (k >= nrun[i]) continue;
      RunDesc d = make_desc(*cs[i], (size_t)(first_run[i] + k), (c128 *)states[i], first);
      first += desc_tiles(d);
      descs.push_back(d);
    }
  }
  starts.push_back(descs.size());
  return launch_run_groups(ctx, descs, starts);
}

}  // namespace qtnh_hip

using namespace qtnh_hip;

extern "C" {

void qtnh_tune_circuit(int tile_bits, int low_bits) {
  g_tile_bits = std::max(3, std::min(tile_bits, kMaxTileBits));
  g_low_bits = std::max(0, std::min(low_bits, 8));
}

void qtnh_tune_circuit_pipeline(int mode) { g_circuit_mode = std::max(0, std::min(mode, 3)); }

qtnh_status qtnh_circuit_create(int n_qubits, qtnh_circuit_t *out) {
  if (!out) return QTNH_EINVAL;
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 62) return QTNH_EINVAL;
  auto *c = new (std::nothrow) qtnh_circuit_s;
  if (!c) return QTNH_ENOMEM;
  c->n = n_qubits;
  *out = c;
  return QTNH_OK;
}

qtnh_status qtnh_circuit_destroy(qtnh_circuit_t c) {
  if (!c) return QTNH_EINVAL;
  if (c->device >= 0) {
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
  }
  free_compiled(*c);
  delete c;
  return QTNH_OK;
}

qtnh_status qtnh_circuit_gate(qtnh_circuit_t c, int k, const uint16_t *qubits, const double *G) {
  if (!c || !qubits || !G || !qubits_valid(c->n, k, qubits)) return QTNH_EINVAL;
  qtnh_circuit_s::Op op;
  op.kind = 0;
  op.k = k;
  for (int j = 0; j < k; ++j) op.q[j] = qubits[j];
  op.coef.assign(G, G + 2 * (1 << k) * (1 << k));
  c->ops.push_back(std::move(op));
  return QTNH_OK;
}

qtnh_status qtnh_circuit_diag(qtnh_circuit_t c, int k, const uint16_t *qubits, const double *d) {
  if (!c || !qubits || !d || !qubits_valid(c->n, k, qubits)) return QTNH_EINVAL;
  qtnh_circuit_s::Op op;
  op.kind = 1;
  op.k = k;
  for (int j = 0; j < k; ++j) op.q[j] = qubits[j];
  op.coef.assign(d, d + 2 * (1 << k));
  c->ops.push_back(std::move(op));
  return QTNH_OK;
}

qtnh_status qtnh_circuit_qft(qtnh_circuit_t c, const double *phases) {
  if (!c) return QTNH_EINVAL;
  const int n = c->n;
  const double h = 0x1.6a09e667f3bcdp-1;
  const double H[8] = {h, 0, h, 0, h, 0, -h, 0};
  for (int i = 0; i < n; ++i) {
    const uint16_t qi = (uint16_t)i;
    qtnh_status st = qtnh_circuit_gate(c, 1, &qi, H);
    if (st) return st;
    for (int j = i + 1; j < n; ++j) {
      const int m = j - i;
      double pr, pi;
      if (phases) {
        pr = phases[2 * m];
        pi = phases[2 * m + 1];
      } else {
        pr = std::cos(M_PI / std::ldexp(1.0, m));
        pi = std::sin(M_PI / std::ldexp(1.0, m));
      }
      const double d[8] = {1, 0, 1, 0, 1, 0, pr, pi};
      const uint16_t q2[2] = {(uint16_t)i, (uint16_t)j};
      st = qtnh_circuit_diag(c, 2, q2, d);
      if (st) return st;
    }
  }
  return QTNH_OK;
}

qtnh_status qtnh_circuit_random(qtnh_circuit_t c, int depth, uint64_t seed) {
  if (!c || depth < 0) return QTNH_EINVAL;
  const int n = c->n;
  HostRng rng(seed);
  std::vector<uint16_t> perm(n);
  for (int layer = 0; layer < depth; ++layer) {
    for (int q = 0; q < n; ++q) {
      double G[8];
      random_su2(rng, G);
      const uint16_t qq = (uint16_t)q;
      qtnh_status st = qtnh_circuit_gate(c, 1, &qq, G);
      if (st) return st;
    }
    for (int q = 0; q < n; ++q) perm[q] = (uint16_t)q;
    for (int q = n - 1; q > 0; --q) std::swap(perm[q], perm[rng.next() % (uint64_t)(q + 1)]);
    const double cz[8] = {1, 0, 1, 0, 1, 0, -1, 0};
    for (int p = 0; p + 1 < n; p += 2) {
      const uint16_t q2[2] = {perm[p], perm[p + 1]};
      qtnh_status st = qtnh_circuit_diag(c, 2, q2, cz);
      if (st) return st;
    }
  }
  return QTNH_OK;
}

qtnh_status qtnh_circuit_grid(qtnh_circuit_t c, int rows, int cols, int depth, uint64_t seed) {
  if (!c || rows < 1 || cols < 1 || depth < 0 || rows * cols != c->n) return QTNH_EINVAL;
  HostRng rng(seed);
  const double cz[8] = {1, 0, 1, 0, 1, 0, -1, 0};
  for (int layer = 0; layer < depth; ++layer) {
    for (int q = 0; q < c->n; ++q) {
      double G[8];
      random_su2(rng, G);
      const uint16_t qq = (uint16_t)q;
      qtnh_status st = qtnh_circuit_gate(c, 1, &qq, G);
      if (st) return st;
    }
    // Couplers cycle through four patterns: horizontal even / odd columns,
    // vertical even / odd rows.
    const int pat = layer % 4;
    for (int r = 0; r < rows; ++r)
      for (int col = 0; col < cols; ++col) {
        int r2 = r, c2 = col;
        if (pat < 2) {
          if ((col & 1) != pat) continue;
          c2 = col + 1;
        } else {
          if ((r & 1) != pat - 2) continue;
          r2 = r + 1;
        }
        if (r2 >= rows || c2 >= cols) continue;
        const uint16_t q2[2] = {(uint16_t)(r * cols + col), (uint16_t)(r2 * cols + c2)};
        qtnh_status st = qtnh_circuit_diag(c, 2, q2, cz);
        if (st) return st;
      }
  }
  return QTNH_OK;
}

qtnh_status qtnh_circuit_info(qtnh_circuit_t c, int *n_qubits, uint64_t *n_ops, uint64_t *n_runs) {
  if (!c) return QTNH_EINVAL;
  if (n_qubits) *n_qubits = c->n;
  if (n_ops) *n_ops = c->ops.size();
  if (n_runs) {
    qtnh_circuit_s tmp;
    tmp.n = c->n;
    tmp.ops = c->ops;
    plan_runs(tmp, c->tuned_tb > 0 ? c->tuned_tb : g_tile_bits, c->tuned_tb > 0 ? c->tuned_low : g_low_bits);
    *n_runs = tmp.runs.size();
  }
  return QTNH_OK;
}

qtnh_status qtnh_circuit_op(qtnh_circuit_t c, uint64_t i, int *kind, int *k, uint16_t *qubits, double *coef) {
  if (!c || i >= c->ops.size()) return QTNH_EINVAL;
  const auto &op = c->ops[i];
  if (kind) *kind = op.kind;
  if (k) *k = op.k;
  if (qubits)
    for (int j = 0; j < op.k; ++j) qubits[j] = op.q[j];
  if (coef) memcpy(coef, op.coef.data(), op.coef.size() * sizeof(double));
  return QTNH_OK;
}

qtnh_status qtnh_circuit_run(qtnh_ctx_t ctx, qtnh_circuit_t c, void *state, int mode) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!c || !state) return set_error(ctx, QTNH_EINVAL, "circuit_run: null argument");
  if (mode == QTNH_RUN_PER_GATE) return run_per_gate(ctx, *c, (c128 *)state);
  if (mode == QTNH_RUN_FUSED) return run_fused(ctx, *c, (c128 *)state);
  return set_error(ctx, QTNH_EINVAL, "circuit_run: unknown mode %d", mode);
}

qtnh_status qtnh_circuit_run_batch(qtnh_ctx_t ctx, int count, const qtnh_circuit_t *cs, void *const *states) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (count < 0 || (count > 0 && (!cs || !states))) return set_error(ctx, QTNH_EINVAL, "run_batch: bad arguments");
  size_t steps = 0;
  for (int i = 0; i < count; ++i) {
    if (!cs[i] || !states[i]) return set_error(ctx, QTNH_EINVAL, "run_batch: null circuit or state %d", i);
    for (int j = 0; j < i; ++j)
      if (states[j] == states[i]) return set_error(ctx, QTNH_EINVAL, "run_batch: state %d repeats state %d", i, j);
    st = compile(ctx, *cs[i]);
    if (st) return st;
    steps = std::max(steps, cs[i]->runs.size());
  }
  // Step s runs the s-th run of every circuit that has one, in one launch.
  std::vector<RunDesc> descs;
  std::vector<size_t> starts;
  for (size_t s = 0; s < steps; ++s) {
    starts.push_back(descs.size());
    uint64_t first = 0;
    for (int i = 0; i < count; ++i) {
      if (s >= cs[i]->runs.size()) continue;
      RunDesc d = make_desc(*cs[i], s, (c128 *)states[i], first);
      first += desc_tiles(d);
      descs.push_back(d);
    }
  }
  starts.push_back(descs.size());
  return launch_run_groups(ctx, descs, starts);
}


qtnh_status qtnh_circuit_tune(qtnh_ctx_t ctx, qtnh_circuit_t c, void *scratch, int *tile_bits, int *low_bits,
                              double *us_out) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!c || !scratch) return set_error(ctx, QTNH_EINVAL, "circuit_tune: null argument");
  if (c->ops.empty()) return QTNH_OK;
  // Candidates: the default tile, then smaller tiles (more workgroups per run;
  // they win for small states whose runs would otherwise occupy a few CUs).
  const int cand[][2] = {{12, 4}, {11, 4}, {10, 2}, {9, 2}, {8, 2}, {7, 2}};
  const int reps = c->n >= 20 ? 2 : 10;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  st = check_hip(ctx, hipEventCreate(&e0), "circuit_tune event");
  if (!st) st = check_hip(ctx, hipEventCreate(&e1), "circuit_tune event");
  double best = 1e30;
  int btb = 0, blow = 0;
  for (const auto &cd : cand) {
    if (st) break;
    if (cd[0] > std::max(c->n, 7)) continue;
    c->tuned_tb = cd[0];
    c->tuned_low = cd[1];
    st = run_fused(ctx, *c, (c128 *)scratch);  // compile + warm up
    float ms = 0;
    if (!st) st = check_hip(ctx, hipEventRecord(e0, ctx->stream), "circuit_tune record");
    for (int r = 0; r < reps && !st; ++r) st = run_fused(ctx, *c, (c128 *)scratch);
    if (!st) st = check_hip(ctx, hipEventRecord(e1, ctx->stream), "circuit_tune record");
    if (!st) st = check_hip(ctx, hipEventSynchronize(e1), "circuit_tune sync");
    if (!st) st = check_hip(ctx, hipEventElapsedTime(&ms, e0, e1), "circuit_tune elapsed");
    if (!st && ms / reps < best) {
      best = ms / reps;
      btb = cd[0];
      blow = cd[1];
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  c->tuned_tb = st ? 0 : btb;
  c->tuned_low = st ? 0 : blow;
  if (st) return st;
  // Leave the winning shape compiled, so the next run launches without a recompile.
  st = compile(ctx, *c);
  if (st) return st;
  if (tile_bits) *tile_bits = btb;
  if (low_bits) *low_bits = blow;
  if (us_out) *us_out = best * 1e3;
  return QTNH_OK;
}

qtnh_status qtnh_autotune_circuit(qtnh_ctx_t ctx, int n_qubits, int *tile_bits, int *low_bits, double *ms_out) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (n_qubits < 1 || n_qubits > 62) return set_error(ctx, QTNH_EINVAL, "autotune: bad qubit count");
  // Proxy: a depth-6 random circuit on min(n, 26) qubits (the best tile does
  // not depend on the state size once it spans many tiles); best of two runs.
  const int np = std::min(n_qubits, 26);
  const uint64_t N = 1ull << np;
  void *w = nullptr;
  st = workspace(ctx, N * sizeof(c128), &w);
  if (st) return st;
  st = launch_fill(ctx, (c128 *)w, N, 1, QTNH_DIST_NORMAL);
  if (st) return st;
  const int cand[][2] = {{10, 3}, {11, 3}, {11, 4}, {12, 4}, {12, 5}, {13, 4}, {13, 5}};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  st = check_hip(ctx, hipEventCreate(&e0), "autotune event");
  if (!st) st = check_hip(ctx, hipEventCreate(&e1), "autotune event");
  double best = 1e30;
  int btb = g_tile_bits, blow = g_low_bits;
  for (const auto &cd : cand) {
    if (st) break;
    if (cd[0] > std::max(np, 10)) continue;
    qtnh_circuit_s c;
    c.n = np;
    (void)qtnh_circuit_random(&c, 6, 12345);
    ctx->circ_tb[np] = (int8_t)cd[0];
    ctx->circ_low[np] = (int8_t)cd[1];
    st = run_fused(ctx, c, (c128 *)w);  // compile + warm up
    float ms = 1e30f;
    for (int rep = 0; rep < 2 && !st; ++rep) {
      float one = 0;
      st = check_hip(ctx, hipEventRecord(e0, ctx->stream), "autotune record");
      if (!st) st = run_fused(ctx, c, (c128 *)w);
      if (!st) st = check_hip(ctx, hipEventRecord(e1, ctx->stream), "autotune record");
      if (!st) st = check_hip(ctx, hipEventSynchronize(e1), "autotune sync");
      if (!st) st = check_hip(ctx, hipEventElapsedTime(&one, e0, e1), "autotune elapsed");
      ms = std::min(ms, one);
    }
    if (!st && ms < best) {
      best = ms;
      btb = cd[0];
      blow = cd[1];
    }
    free_compiled(c);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) {
    ctx->circ_tb[np] = 0;
    return st;
  }
  ctx->circ_tb[n_qubits] = (int8_t)btb;
  ctx->circ_low[n_qubits] = (int8_t)blow;
  if (np != n_qubits) {
    ctx->circ_tb[np] = (int8_t)btb;
    ctx->circ_low[np] = (int8_t)blow;
  }
  if (tile_bits) *tile_bits = btb;
  if (low_bits) *low_bits = blow;
  if (ms_out) *ms_out = best;
  return QTNH_OK;
}

qtnh_status qtnh_expect_z_c128(qtnh_ctx_t ctx, const void *state, int n_qubits, int nq, const uint16_t *qubits,
                               double *out) {
  qtnh_status st = enter_ctx(ctx);
  if (st) return st;
  if (!state || !out || n_qubits < 1 || n_qubits > 62 || nq < 1 || nq > 1024 || !qubits)
    return set_error(ctx, QTNH_EINVAL, "expect_z: bad arguments");
  for (int j = 0; j < nq; ++j)
    if (qubits[j] >= n_qubits) return set_error(ctx, QTNH_EINVAL, "expect_z: qubit %d out of range", qubits[j]);
  const uint64_t n = 1ull << n_qubits;
  const int cb = std::min(n_qubits, kChunkBits);
  const uint64_t nchunks = n >> cb;
  const uint64_t part_bytes = ((nchunks * kZSlots * sizeof(double)) + 255) & ~255ull;
  st = grow_buffer(ctx, &ctx->red, &ctx->red_bytes, part_bytes + (uint64_t)nq * (sizeof(int) + sizeof(double)) + 512);
  if (st) return st;
  double *part = (double *)ctx->red;
  int *bits = (int *)((char *)ctx->red + part_bytes);
  double *dout = (double *)((char *)ctx->red + part_bytes + (((uint64_t)nq * sizeof(int) + 255) & ~255ull));
  std::vector<int> hb(nq);
  for (int j = 0; j < nq; ++j) hb[j] = n_qubits - 1 - qubits[j];
  st = check_hip(ctx, hipMemcpyAsync(bits, hb.data(), nq * sizeof(int), hipMemcpyHostToDevice, ctx->stream),
                 "expect_z upload");
  if (st) return st;
  const unsigned blocks = (unsigned)std::min<uint64_t>(nchunks, (uint64_t)ctx->num_cus * 8);
  hipLaunchKernelGGL(zchunk_kernel, dim3(blocks), dim3(kZThreads), 0, ctx->stream, (const double2_t *)state, n, cb,
                     part, nchunks);
  st = check_hip(ctx, hipGetLastError(), "expect_z launch");
  if (st) return st;
  hipLaunchKernelGGL(zfinal_kernel, dim3(nq), dim3(kZThreads), 0, ctx->stream, part, nchunks, cb, bits, dout);
  st = check_hip(ctx, hipGetLastError(), "expect_z final launch");
  if (st) return st;
  st = check_hip(ctx, hipMemcpyAsync(out, dout, nq * sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
                 "expect_z copy");
  if (st) return st;
  return check_hip(ctx, hipStreamSynchronize(ctx->stream), "expect_z sync");
}

}  // extern "C"
