// Internal declarations shared by the HIP translation units of libqtnh_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "qtnh_hip.h"

struct qtnh_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;  // caller's stream (not owned)
  int num_cus = 0;
  std::string last_error;
  // Growable device workspace for permute-to-matrix temporaries.
  void *ws = nullptr;
  uint64_t ws_bytes = 0;
  // Small pinned host buffer + device buffer for reduction results.
  double *host_scalar = nullptr;
  double *dev_scalar = nullptr;
  // Growable device buffer for launch descriptors (circuit runs, batches).
  void *desc = nullptr;
  uint64_t desc_bytes = 0;
  // Growable device buffer for reduction partials (expectation values).
  void *red = nullptr;
  uint64_t red_bytes = 0;
  // Growable pinned host buffer staging descriptor uploads (asynchronous copies).
  void *pinned = nullptr;
  uint64_t pinned_bytes = 0;
  // Second stream (default priority) for work that overlaps the caller's
  // stream inside one call, forked and joined with the two events.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Fused-circuit tile shape chosen by the autotuner, per qubit count (0: unset).
  int8_t circ_tb[64] = {0};
  int8_t circ_low[64] = {0};
  // Set for the duration of a call that must create no HIP objects and not
  // synchronise (qtnh_contract_c128_noalloc): kernels that time their tile
  // shapes on first use take the cached or the default shape instead.
  bool fixed_shapes = false;
};

namespace qtnh_hip {

// Complex element as stored in HBM: 16 bytes, (re, im).
struct c128 {
  double re, im;
};

constexpr int kMaxRank = QTNH_MAX_RANK;

qtnh_status set_error(qtnh_ctx_t ctx, qtnh_status st, const char *fmt, ...);
qtnh_status check_hip(qtnh_ctx_t ctx, hipError_t e, const char *what);
// Ensures ctx->ws holds at least `bytes`; returns its pointer in *out.
qtnh_status workspace(qtnh_ctx_t ctx, uint64_t bytes, void **out);
// Grows a context-owned device buffer (*buf, *cap) to at least `bytes`; the old
// buffer is freed only after the stream has drained.
qtnh_status grow_buffer(qtnh_ctx_t ctx, void **buf, uint64_t *cap, uint64_t bytes);
// Grows the context's pinned host staging buffer to at least `bytes`; the old
// buffer is freed only after the stream has drained (queued copies read it).
qtnh_status grow_pinned(qtnh_ctx_t ctx, uint64_t bytes);
// Makes the context's device current for the calling thread.
qtnh_status enter_ctx(qtnh_ctx_t ctx);

// Complex product without FMA contraction (files are built with -ffp-contract=off):
// the arithmetic of std::complex<double> operator* in a GCC build without -march.
__host__ __device__ inline void cmul_parts(double ar, double ai, double br, double bi, double *r, double *i) {
  *r = ar * br - ai * bi;
  *i = ar * bi + ai * br;
}

// Workgroup barrier ordering LDS accesses only: a wave waits for its own LDS
// operations (lgkmcnt), not for global loads, stores or LDS DMA still in
// flight on the vector-memory counter.  The "memory" clobbers keep the
// compiler from moving memory accesses across it.  (__syncthreads' workgroup
// fence also waits for those, and for the DMA, which writes LDS.)
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Row-major size with overflow detection; false on overflow.
bool checked_size(int rank, const uint64_t *dims, uint64_t *n);

// Binary view of a tensor whose dims are all powers of two: axis i occupies
// index bits [shift[i], shift[i] + width[i]) (row-major: later axes lower).
struct BitLayout {
  bool binary = false;  // every dim is a power of two
  int nbits = 0;
  int shift[kMaxRank];
  int width[kMaxRank];
};
BitLayout bit_layout(int rank, const uint64_t *dims);

// ---- kernels (launch wrappers; argument checks happen in the ABI layer) ----
qtnh_status launch_fill(qtnh_ctx_t ctx, c128 *buf, uint64_t n, uint64_t seed, int dist);
qtnh_status launch_permute(qtnh_ctx_t ctx, const c128 *src, c128 *dst, int rank, const uint64_t *dims,
                           const uint16_t *ptup);
// A square dense gate of gate_bits (1 to 3) term bits contracted into A with
// its outputs on the lowest result bits, as the bit permutation sigma (A bit
// -> C bit, the term bits to C bits 0 .. gate_bits-1) with each result row
// formed in the tiles (reference term order and arithmetic); xb[t 2^gb + o]:
// B offset of term t for output o.  nb: A's index bits.  *done false when the
// shape does not fit the tiles (nothing launched; the caller takes another
// kernel).  C must not overlap A.
qtnh_status launch_permute_gate(qtnh_ctx_t ctx, const c128 *A, c128 *C, const c128 *B, int nb, const int *sigma,
                                int gate_bits, const uint32_t *xb, bool *done);
qtnh_status launch_swap(qtnh_ctx_t ctx, c128 *buf, int rank, const uint64_t *dims, int i, int j);
qtnh_status launch_gate(qtnh_ctx_t ctx, c128 *state, int rank, const uint64_t *dims, int k,
                        const uint16_t *axes, const double *G);
qtnh_status launch_diag(qtnh_ctx_t ctx, c128 *state, int rank, const uint64_t *dims, int k,
                        const uint16_t *axes, const double *d);
// ext_ws (optional): a workspace region used instead of the context's growable one.
qtnh_status launch_contract(qtnh_ctx_t ctx, const c128 *A, int rA, const uint64_t *dA, const c128 *B, int rB,
                            const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *repl1,
                            const uint16_t *repl2, c128 *C, int precision, double *err_out, void *ext_ws = nullptr,
                            uint64_t ext_ws_bytes = 0);
// Route launch_contract takes for these arguments, and its workspace need.
struct ContractRoute {
  bool direct;
  uint64_t M, N, K, ws_bytes;
};
ContractRoute contract_route(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                             const uint16_t *repl1, const uint16_t *repl2, int precision);
// The GEMM route of launch_contract for these arguments (false: the direct
// route): A' = A permuted by pA to [open1, closed]; B' = B permuted by pB to
// [closed, open2] (FP64 GEMM) or [open2, closed] (split-K, split precision);
// the natural result [open1, open2] (dims dn) permuted by pC to the placement.
// An operand already in its layout (idA / idB / idC) is used in place.
// Workspace (bytes from its start): A' at offA, B' at offB, natural C at
// offC, then split scratch, split-K partials or depth-chunk partials at offS.
struct GemmRoute {
  bool splitk = false, split = false;
  bool b_cols = true;  // B' is [closed, open2] (K x N); else [open2, closed]
  bool idA = true, idB = true, idC = true;
  std::vector<uint16_t> pA, pB, pC;
  std::vector<uint64_t> dn;
  uint64_t M = 0, N = 0, K = 0, nC = 0;
  uint64_t kchunks = 0, zchunks = 1, sbytes = 0, bytes = 0;
  uint64_t offA = 0, offB = 0, offC = 0, offS = 0;
};
bool gemm_route(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                const uint16_t *repl1, const uint16_t *repl2, int precision, GemmRoute *out);
// Direct kernels only (reference summation order, no workspace use).
qtnh_status launch_contract_direct(qtnh_ctx_t ctx, const c128 *A, int rA, const uint64_t *dA, const c128 *B, int rB,
                                   const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *repl1,
                                   const uint16_t *repl2, c128 *C);
// Argument checks of qtnh_contract_c128 and its no-allocation variant; *nC
// receives the result's element count (0: nothing to do).
qtnh_status check_contract_args(qtnh_ctx_t ctx, const void *A, int rA, const uint64_t *dA, const void *B, int rB,
                                const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *repl1,
                                const uint16_t *repl2, const void *C, int precision, uint64_t *nC);
// Looks up the device address of the no-allocation scratch array on the
// context's device (current), once per device (noalloc.hip).
qtnh_status noalloc_setup(qtnh_ctx_t ctx);
qtnh_status launch_norm2(qtnh_ctx_t ctx, const c128 *x, uint64_t n, double *out);
qtnh_status launch_scale(qtnh_ctx_t ctx, c128 *x, uint64_t n, double re, double im);
// Zero fill with a kernel (capturable in graphs, unlike a memset node).
qtnh_status launch_zero(qtnh_ctx_t ctx, void *p, uint64_t bytes);

// Complex GEMM C[MxN] = A[MxK] * B[KxN], row-major with leading dims lda/ldb/ldc (elements).
qtnh_status launch_zgemm_f64(qtnh_ctx_t ctx, const c128 *A, const c128 *B, c128 *C, uint64_t M, uint64_t N,
                             uint64_t K, uint64_t lda, uint64_t ldb, uint64_t ldc, c128 *partials);
// Depth chunks launch_zgemm_f64 uses for this shape (1: none); with more than
// one it needs `partials` of chunks * M * N elements.
uint64_t zgemm_chunks(uint64_t M, uint64_t N, uint64_t K);

// C[MxN] = A[MxK] * Bt[NxK]^T for few outputs: each sum in `chunks` ordered
// chunks over separate workgroups (partials: M*N*chunks elements), then summed
// in chunk order.
qtnh_status launch_zgemm_splitk(qtnh_ctx_t ctx, const c128 *A, const c128 *Bt, c128 *C, uint64_t M, uint64_t N,
                                uint64_t K, uint64_t chunks, c128 *partials);

// ---- grouped launches: many independent problems in one launch, each a
// contiguous range of blocks starting at `first` (descriptors sorted by it).
// One GEMM of launch_zgemm_grouped: C[MxN] = A[MxK] * B[KxN], all contiguous
// row-major; `tiles` output tiles (64 x block width) times the depth chunks of
// zgemm_chunk_shape (chunk c sums k in [c kc, (c + 1) kc) into C + c cstride).
struct ZGroupDesc {
  const c128 *A;
  const c128 *B;
  c128 *C;
  uint64_t M, N, K, kc, cstride;
  uint64_t first, tiles;
};
// Depth-chunk sums: C[i] = sum over c of part[c MN + i] (256 outputs per block).
struct ZSumDesc {
  const c128 *part;
  c128 *C;
  uint64_t MN, chunks, first;
};
// Split-K contractions: partial blocks [first, first + MN S), final blocks
// [first2, first2 + ceil(MN / 256)).
struct SkDesc {
  const c128 *A;
  const c128 *Bt;
  c128 *part;
  c128 *C;
  uint64_t N, K, kc, S, MN;
  uint64_t first, first2;
};
// A bit permutation (every dim a power of two): element j of dst is element
// sum_p bit_p(j) << sbit[p] of src; 1024 elements per block.
struct PermDesc {
  const c128 *src;
  c128 *dst;
  uint64_t n, first;
  int nb;
  int8_t sbit[64];
};
qtnh_status launch_zgemm_grouped(qtnh_ctx_t ctx, int bn, const ZGroupDesc *d, int nd, uint64_t total_blocks);
qtnh_status launch_zgemm_grouped_sum(qtnh_ctx_t ctx, const ZSumDesc *d, int nd, uint64_t total_blocks);
qtnh_status launch_splitk_grouped(qtnh_ctx_t ctx, const SkDesc *d, int nd, uint64_t total_partial,
                                  uint64_t total_final);
qtnh_status launch_permute_grouped(qtnh_ctx_t ctx, const PermDesc *d, int nd, uint64_t total_blocks);
// Block width (64 or 128) and depth chunks (length on BK steps, count) of launch_zgemm_f64.
int zgemm_block_n(uint64_t M, uint64_t N);
void zgemm_chunk_shape(uint64_t M, uint64_t N, uint64_t K, uint64_t *kc, uint64_t *nchunks);
// Terms per chunk of the split-K route (a multiple of its workgroup size).
uint64_t splitk_chunk_len(uint64_t K, uint64_t chunks);
// Source bit of every result bit of a permutation of power-of-two dims
// (false when a dim is not a power of two).
bool permute_source_bits(int rank, const uint64_t *dims, const uint16_t *ptup, int *nb, int8_t *sbit);

// Split-precision complex GEMM C[MxN] = A[MxK] * B (all contiguous, row-major)
// on INT8 matrix cores, B given as B[KxN] (b_cols) or as Bt[NxK]; scratch of
// split_scratch_bytes(M, N, K).
qtnh_status launch_zgemm_split(qtnh_ctx_t ctx, const c128 *A, const c128 *B, bool b_cols, c128 *C, uint64_t M,
                               uint64_t N, uint64_t K, void *scratch, uint64_t scratch_bytes, double *err_out);
uint64_t split_scratch_bytes(uint64_t M, uint64_t N, uint64_t K);
int split_slices();

// Fused-circuit pieces used by the serving loop.  circuit_run_count compiles
// the circuit for this context and returns its number of fused runs;
// circuit_advance applies runs first_run[i] .. first_run[i] + nrun[i] - 1 of
// circuit i to state i, the k-th of every circuit in one launch (k = 0, 1, ..).
qtnh_status circuit_run_count(qtnh_ctx_t ctx, qtnh_circuit_t c, int *nruns);
qtnh_status circuit_advance(qtnh_ctx_t ctx, int count, const qtnh_circuit_t *cs, void *const *states,
                            const int *first_run, const int *nrun);

// Host helpers for placement (also exported through the ABI).
qtnh_status default_repls(int r1, int ndis1, int r2, int ndis2, int nw, const uint16_t *wires, uint16_t *repl1,
                          uint16_t *repl2);
qtnh_status symm_repls(int r1, int nd2, int nl2, int n_dis_in, int nw, const uint16_t *wires, uint16_t *repl1,
                       uint16_t *repl2, int *applies);
qtnh_status contract_dims(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                          const uint16_t *repl1, const uint16_t *repl2, int *rC, uint64_t *dC);

// Grid helper: at most `cap` blocks of `threads`, enough for `work` items.
inline unsigned grid_for(uint64_t work, unsigned threads, unsigned cap) {
  uint64_t b = (work + threads - 1) / threads;
  if (b == 0) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

}  // namespace qtnh_hip
