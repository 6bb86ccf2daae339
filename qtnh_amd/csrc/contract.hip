// General pairwise contraction (reference src/tensor/contract.cpp:12-201).
//
// Two routes:
//  * direct: one thread per result element, the closed tuples summed in the
//    reference order (wires sorted by their t2 axis, last one fastest,
//    accumulator from 0, products without FMA; contract.cpp:26, 131-156 and
//    indexing.cpp:27-35).  Used when the contraction is small or one operand
//    is gate-sized; it reproduces the reference's bits.
//  * GEMM: permute A to [open1, closed] and B to [closed, open2], complex GEMM
//    on MFMA, then permute the natural result [open1, open2] to the
//    placement.  Used when both sides are large (within 1e-12 relative).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "internal.hpp"

namespace qtnh_hip {
namespace {

constexpr int kDirThreads = 256;
constexpr int kMaxOut = 48;     // result axes handled by the direct kernels
constexpr int kMaxClosed = 24;  // closed axes handled by the direct kernels

__device__ inline void cmac(double &ar, double &ai, c128 a, c128 b) {
  double pr = a.re * b.re - a.im * b.im;
  double pi = a.re * b.im + a.im * b.re;
  ar = ar + pr;
  ai = ai + pi;
}

struct Closed {
  uint64_t kd[kMaxClosed];  // dim of closed tuple position w (reference order)
  uint64_t ka[kMaxClosed];  // A stride
  uint64_t kb[kMaxClosed];  // B stride
  int nw;
};

// Closed tuples of a contraction whose dims are all powers of two: tuple t's
// bit c moves the A / B offsets by a power of two, so counting t up by one
// flips the offset bits in the prefix masks fa / fb[ctz(t + 1)].
constexpr int kMaxClosedBits = 32;
struct ClosedBits {
  uint64_t fa[kMaxClosedBits];
  uint64_t fb[kMaxClosedBits];
  int kbits;
};

// Sum over the closed tuples in reference order (odometer, last position fastest).
__device__ inline c128 closed_sum(const c128 *__restrict__ A, const c128 *__restrict__ B, uint64_t oa, uint64_t ob,
                                  const Closed &cl) {
  double ar = 0.0, ai = 0.0;
  uint32_t idx[kMaxClosed];
  for (int w = 0; w < cl.nw; ++w) idx[w] = 0;
  while (true) {
    cmac(ar, ai, A[oa], B[ob]);
    int w = cl.nw - 1;
    for (; w >= 0; --w) {
      if (++idx[w] < cl.kd[w]) {
        oa += cl.ka[w];
        ob += cl.kb[w];
        break;
      }
      oa -= (cl.kd[w] - 1) * cl.ka[w];
      ob -= (cl.kd[w] - 1) * cl.kb[w];
      idx[w] = 0;
    }
    if (w < 0) break;
  }
  c128 r;
  r.re = ar;
  r.im = ai;
  return r;
}

// Result axes in C memory order: dim, and the A / B stride of the open axis placed there.
struct OutGen {
  uint64_t cd[kMaxOut];
  uint64_t ca[kMaxOut];
  uint64_t cb[kMaxOut];
  int rc;
};

__global__ void __launch_bounds__(kDirThreads) contract_direct_general(const c128 *__restrict__ A,
                                                                       const c128 *__restrict__ B,
                                                                       c128 *__restrict__ C, OutGen out, Closed cl,
                                                                       uint64_t nC) {
  for (uint64_t e = (uint64_t)blockIdx.x * kDirThreads + threadIdx.x; e < nC; e += (uint64_t)gridDim.x * kDirThreads) {
    uint64_t rest = e, oa = 0, ob = 0;
    for (int p = out.rc - 1; p >= 0; --p) {
      uint64_t d = out.cd[p];
      uint64_t i = rest % d;
      rest /= d;
      oa += i * out.ca[p];
      ob += i * out.cb[p];
    }
    C[e] = closed_sum(A, B, oa, ob, cl);
  }
}

// Power-of-two result dims: C index bit p comes from A bit abit[p] or B bit
// bbit[p] (-1 when from the other side).  A block handles 256 consecutive
// result elements per step; the low 8 bits are per-thread constants.
struct OutBits {
  int8_t abit[64];
  int8_t bbit[64];
  int nbits;
};

__global__ void __launch_bounds__(kDirThreads) contract_direct_bits(const c128 *__restrict__ A,
                                                                    const c128 *__restrict__ B, c128 *__restrict__ C,
                                                                    OutBits out, ClosedBits cl, uint64_t nC) {
  const int lowbits = out.nbits < 8 ? out.nbits : 8;
  __shared__ uint64_t flip_a[kMaxClosedBits], flip_b[kMaxClosedBits];
  if ((int)threadIdx.x < cl.kbits) {
    flip_a[threadIdx.x] = cl.fa[threadIdx.x];
    flip_b[threadIdx.x] = cl.fb[threadIdx.x];
  }
  // A and B offsets of every value of each group of 8 high result bits, so
  // that an element's offsets are a few table reads ORed together.
  constexpr int kTabs = 7;  // 8 low + 7 x 8 high bits cover 64
  __shared__ uint64_t tab_a[kTabs][256], tab_b[kTabs][256];
  const int ntab = (out.nbits - lowbits + 7) / 8;
  for (int t = 0; t < ntab; ++t) {
    uint64_t va = 0, vb = 0;
    for (int j = 0; j < 8; ++j) {
      const int p = lowbits + 8 * t + j;
      if (p >= out.nbits || !((threadIdx.x >> j) & 1u)) continue;
      if (out.abit[p] >= 0) va |= 1ull << out.abit[p];
      if (out.bbit[p] >= 0) vb |= 1ull << out.bbit[p];
    }
    tab_a[t][threadIdx.x] = va;
    tab_b[t][threadIdx.x] = vb;
  }
  __syncthreads();
  uint64_t la = 0, lb = 0;
  for (int p = 0; p < lowbits; ++p) {
    uint64_t bit = (threadIdx.x >> p) & 1u;
    if (out.abit[p] >= 0) la |= bit << out.abit[p];
    if (out.bbit[p] >= 0) lb |= bit << out.bbit[p];
  }
  const uint64_t nhi = nC >> lowbits;
  if (threadIdx.x >= (1u << lowbits)) return;
  for (uint64_t h = blockIdx.x; h < nhi; h += gridDim.x) {
    uint64_t ha = la, hb = lb;
    for (int t = 0; t < ntab; ++t) {
      const uint32_t v = (uint32_t)(h >> (8 * t)) & 255u;
      ha |= tab_a[t][v];
      hb |= tab_b[t][v];
    }
    // Terms in reference order; the loads of four terms are issued before
    // their products are accumulated (same products, same order).
    const uint64_t K = 1ull << cl.kbits;
    double ar = 0.0, ai = 0.0;
    uint64_t xa = ha, xb = hb;
    for (uint64_t t = 0; t < K; t += 4) {
      c128 a[4], b[4];
      const int un = K - t < 4 ? (int)(K - t) : 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (u >= un) break;
        a[u] = A[xa];
        b[u] = B[xb];
        const uint64_t tn = t + (uint64_t)u + 1;
        if (tn < K) {
          const int c = __builtin_ctzll(tn);
          xa ^= flip_a[c];
          xb ^= flip_b[c];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (u >= un) break;
        cmac(ar, ai, a[u], b[u]);
      }
    }
    c128 r;
    r.re = ar;
    r.im = ai;
    C[(h << lowbits) | threadIdx.x] = r;
  }
}

// Power-of-two dims with few open B results on the lowest result bits (a
// small gate contracted into a large state with its outputs placed last): one
// thread per row of N = 2^NB results that share their A elements.  Each
// result is 0 + the sum over its K = 2^KB terms in reference order (the order
// and arithmetic of contract_direct_bits).
//
// A block takes 256 R consecutive rows per step (R = 2^RB rows per thread).
// Its A elements (256 R rows x K terms) are read in the order of the A bits
// they vary in, lowest first, so a wave's loads cover contiguous bytes; they
// pass through LDS to the thread of their row (K <= 8).  The results pass
// through LDS too, up to 8 of each row at a time, and are written back so
// that consecutive lanes store consecutive elements (whole rows, or 128-byte
// halves of 16-result rows).  LDS slots are padded by one per 16 so both
// sides of each exchange are free of bank conflicts.  nrows is a multiple of
// 256 R (at least 12 row bits, RB <= 1), so every thread of a block runs the
// same steps.
//
// The barriers between the exchanges order LDS accesses only: the next
// step's A loads, issued before this step's results are formed, and this
// step's C stores stay in flight across them.  The first step is peeled so
// that the loop is entered with the same loads and stores outstanding as on
// its back edge (the compiler's wait before the staged values are written to
// LDS then covers the loads only, not the stores issued after them).  A and C
// are not restrict-qualified so that the compiler keeps the loads ahead of
// the barriers and stores.
template <int NB, int KB>
struct RowsShape {
  static constexpr int N = 1 << NB, K = 1 << KB;
  static constexpr bool kStageIn = KB <= 3;
  static constexpr int NS = N > 8 ? 8 : N;  // results of a row staged per pass
  // Two rows per thread where both exchanges fit 2048 LDS slots (32 KiB):
  // twice the loads in flight per thread.
  static constexpr int RB = kStageIn && (K > NS ? K : NS) <= 4 ? 1 : 0;
  static constexpr int R = 1 << RB;
  static constexpr int kRows = kDirThreads * R;  // rows per block step
  static constexpr int kSlots = (kStageIn && K > NS ? K : NS) * kRows;
  static constexpr int kStaged = kStageIn ? K * R : 1;  // staged A elements per thread
};

template <int NB, int KB>
struct RowsThread : RowsShape<NB, KB> {
  using S = RowsShape<NB, KB>;
  // Staged reads: element j of this thread is number e = 256 j + tid of the
  // block's 256 R K A elements in the order of the A bits they vary in
  // (ascending); off[j] is its A offset within the block, slot[j] its LDS slot.
  uint64_t off[S::kStaged];
  int slot[S::kStaged];
  // Unstaged reads: A offset of term t within a row.
  uint64_t da[S::kStageIn ? 1 : S::K];
  // Staged A values in flight (real and imaginary parts apart, so they stay
  // in registers from one step to the next).
  double vre[S::kStaged], vim[S::kStaged];
};

// A offset of the block step starting at row r (its low 8 + RB row bits are zero).
__device__ __forceinline__ uint64_t rows_base(const uint64_t (*tab_a)[256], int ntab, uint64_t r) {
  uint64_t base = 0;
  for (int t = 1; t < ntab; ++t) base |= tab_a[t][(uint32_t)(r >> (8 * t)) & 255u];
  return base;
}

template <int NB, int KB>
__device__ __forceinline__ void rows_load(RowsThread<NB, KB> &T, const c128 *A, uint64_t base) {
#pragma unroll
  for (int j = 0; j < RowsThread<NB, KB>::kStaged; ++j) {
    const c128 x = A[base | T.off[j]];
    T.vre[j] = x.re;
    T.vim[j] = x.im;
  }
}

// One step: the block's rows r0 .. r0 + 256 R - 1 (this thread's: r0 + tid + 256 q).
template <int NB, int KB>
__device__ __forceinline__ void rows_step(RowsThread<NB, KB> &T, const c128 *A, c128 *C,
                                          c128 *stage, const uint64_t (*tab_a)[256], const c128 *bco, int ntab,
                                          uint64_t r0, uint64_t stride, uint64_t nrows, int tid) {
  using S = RowsShape<NB, KB>;
  constexpr int N = S::N, K = S::K, NS = S::NS, R = S::R, kRows = S::kRows;
  c128 a[R][K];
  if constexpr (S::kStageIn) {
#pragma unroll
    for (int j = 0; j < S::kStaged; ++j) {
      stage[T.slot[j]].re = T.vre[j];
      stage[T.slot[j]].im = T.vim[j];
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int t = 0; t < K; ++t) {
        const int sl = t * kRows + q * kDirThreads + tid;
        a[q][t] = stage[sl + (sl >> 4)];
      }
    // Next step's loads; the last step reloads its own block (in bounds, unused).
    const uint64_t r1 = r0 + stride < nrows ? r0 + stride : r0;
    rows_load(T, A, rows_base(tab_a, ntab, r1));
    lds_barrier();
  } else {
    const uint64_t oa = rows_base(tab_a, ntab, r0) | tab_a[0][tid];
#pragma unroll
    for (int t = 0; t < K; ++t) a[0][t] = A[oa | T.da[t]];
  }
#pragma unroll
  for (int h = 0; h < N / NS; ++h) {
#pragma unroll
    for (int q = 0; q < R; ++q)
#pragma unroll
      for (int p = 0; p < NS; ++p) {
        const int o = h * NS + p;
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int t = 0; t < K; ++t) cmac(ar, ai, a[q][t], bco[t * N + o]);
        c128 v;
        v.re = ar;
        v.im = ai;
        if constexpr (N == 1) {
          C[r0 + q * kDirThreads + tid] = v;
        } else {
          const int sl = (q * kDirThreads + tid) * NS + p;
          stage[sl + (sl >> 4)] = v;
        }
      }
    if constexpr (N > 1) {
      lds_barrier();
#pragma unroll
      for (int j = 0; j < NS * R; ++j) {
        const int e = j * kDirThreads + tid;
        C[(r0 + e / NS) * N + h * NS + e % NS] = stage[e + (e >> 4)];
      }
      lds_barrier();
    }
  }
}

template <int NB, int KB>
__global__ void __launch_bounds__(kDirThreads) contract_direct_rows(const c128 *A, const c128 *__restrict__ B, c128 *C,
                                                                    OutBits out, ClosedBits cl, uint64_t nrows) {
  using S = RowsShape<NB, KB>;
  constexpr int N = S::N, K = S::K, RB = S::RB;
  constexpr int kTabs = 7;
  __shared__ uint64_t tab_a[kTabs][256];
  __shared__ c128 bco[K * N];  // bco[t N + o]: B element of term t for result o
  __shared__ c128 stage[S::kSlots + S::kSlots / 16];
  const int tid = threadIdx.x;
  const int rbits = out.nbits - NB;
  const int ntab = (rbits + 7) / 8;
  for (int t = 0; t < ntab; ++t) {
    uint64_t va = 0;
    for (int j = 0; j < 8; ++j) {
      const int p = NB + 8 * t + j;
      if (p < out.nbits && ((tid >> j) & 1)) va |= 1ull << out.abit[p];
    }
    tab_a[t][tid] = va;
  }
  for (int i = tid; i < K * N; i += kDirThreads) {
    const int t = i / N, o = i % N;
    uint64_t xb = 0;
    for (int c = 0; c < KB; ++c)
      if ((t >> c) & 1) xb |= cl.fb[c] ^ (c ? cl.fb[c - 1] : 0ull);
    for (int j = 0; j < NB; ++j)
      if ((o >> j) & 1) xb |= 1ull << out.bbit[j];
    bco[i] = B[xb];
  }
  // A bit of each term bit c.
  int tbit[KB];
#pragma unroll
  for (int c = 0; c < KB; ++c) tbit[c] = __builtin_ctzll(cl.fa[c] ^ (c ? cl.fa[c - 1] : 0ull));
  RowsThread<NB, KB> T;
  if constexpr (S::kStageIn) {
    // Varying bits as (A bit << 8) | w, w = row bit j (< 16) or 16 + term bit c,
    // sorted by A bit.
    constexpr int kVary = 8 + RB + KB;
    int key[kVary];
#pragma unroll
    for (int j = 0; j < 8 + RB; ++j) key[j] = ((int)out.abit[NB + j] << 8) | j;
#pragma unroll
    for (int c = 0; c < KB; ++c) key[8 + RB + c] = (tbit[c] << 8) | (16 + c);
#pragma unroll
    for (int i = 1; i < kVary; ++i)
#pragma unroll
      for (int m = i; m > 0; --m)
        if (key[m - 1] > key[m]) {
          const int x = key[m - 1];
          key[m - 1] = key[m];
          key[m] = x;
        }
#pragma unroll
    for (int j = 0; j < S::kStaged; ++j) {
      const int e = j * kDirThreads + tid;
      uint64_t o = 0;
      int row = 0, t = 0;
#pragma unroll
      for (int i = 0; i < kVary; ++i) {
        if (!((e >> i) & 1)) continue;
        o |= 1ull << (key[i] >> 8);
        const int w = key[i] & 255;
        if (w < 16)
          row |= 1 << w;
        else
          t |= 1 << (w - 16);
      }
      T.off[j] = o;
      const int sl = t * S::kRows + row;  // term-major
      T.slot[j] = sl + (sl >> 4);
    }
  } else {
#pragma unroll
    for (int t = 0; t < K; ++t) {
      uint64_t x = 0;
#pragma unroll
      for (int c = 0; c < KB; ++c)
        if ((t >> c) & 1) x |= 1ull << tbit[c];
      T.da[t] = x;
    }
  }
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * S::kRows;
  uint64_t r0 = (uint64_t)blockIdx.x * S::kRows;
  if (r0 >= nrows) return;
  if constexpr (S::kStageIn) rows_load(T, A, rows_base(tab_a, ntab, r0));
  rows_step(T, A, C, stage, tab_a, bco, ntab, r0, stride, nrows, tid);
  for (r0 += stride; r0 < nrows; r0 += stride) rows_step(T, A, C, stage, tab_a, bco, ntab, r0, stride, nrows, tid);
}

// Row kernels exist for NB + KB <= 5 (at most 32 coefficients, held in registers).
constexpr int kRowsMaxBits = 5;
template <int NB, int KB>
void launch_rows_kb(int kb, dim3 g, hipStream_t st, const c128 *A, const c128 *B, c128 *C, const OutBits &ob,
                    const ClosedBits &cb, uint64_t nrows) {
  if constexpr (KB >= 1 && NB + KB <= kRowsMaxBits) {
    if (kb == KB) {
      const uint64_t steps = nrows / RowsShape<NB, KB>::kRows;
      const dim3 gk((unsigned)std::min<uint64_t>(std::max<uint64_t>(steps, 1), g.x));
      hipLaunchKernelGGL((contract_direct_rows<NB, KB>), gk, dim3(kDirThreads), 0, st, A, B, C, ob, cb, nrows);
      return;
    }
    launch_rows_kb<NB, KB - 1>(kb, g, st, A, B, C, ob, cb, nrows);
  } else if constexpr (KB >= 1) {
    launch_rows_kb<NB, KB - 1>(kb, g, st, A, B, C, ob, cb, nrows);
  }
}

void strides(int r, const uint64_t *d, uint64_t *st) {
  uint64_t s = 1;
  for (int i = r - 1; i >= 0; --i) {
    st[i] = s;
    s *= d[i];
  }
}

// Everything the routes need about one contraction.
struct Plan {
  int rA, rB, rC, nw;
  uint64_t dC[kMaxRank];
  std::vector<int> open1, open2;   // ascending axis lists
  std::vector<int> cA, cB;         // closed axes in reference order
  uint64_t M = 1, N = 1, K = 1, nC = 1;
  std::vector<int> pos;            // result position of natural axis q
};

Plan make_plan(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
               const uint16_t *r1, const uint16_t *r2) {
  Plan P;
  P.rA = rA;
  P.rB = rB;
  P.nw = nw;
  contract_dims(rA, dA, rB, dB, nw, wires, r1, r2, &P.rC, P.dC);
  std::vector<int> order(nw);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int x, int y) { return wires[2 * x + 1] < wires[2 * y + 1]; });
  std::vector<char> wa(rA, 0), wb(rB, 0);
  for (int w : order) {
    P.cA.push_back(wires[2 * w]);
    P.cB.push_back(wires[2 * w + 1]);
    wa[wires[2 * w]] = 1;
    wb[wires[2 * w + 1]] = 1;
    P.K *= dA[wires[2 * w]];
  }
  for (int i = 0; i < rA; ++i)
    if (!wa[i]) {
      P.open1.push_back(i);
      P.M *= dA[i];
      P.pos.push_back(r1[i]);
    }
  for (int i = 0; i < rB; ++i)
    if (!wb[i]) {
      P.open2.push_back(i);
      P.N *= dB[i];
      P.pos.push_back(r2[i]);
    }
  P.nC = P.M * P.N;
  return P;
}

}  // namespace

qtnh_status launch_contract_direct(qtnh_ctx_t ctx, const c128 *A, int rA, const uint64_t *dA, const c128 *B, int rB,
                                   const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *r1,
                                   const uint16_t *r2, c128 *C) {
  Plan P = make_plan(rA, dA, rB, dB, nw, wires, r1, r2);
  if (P.nw > kMaxClosed || P.rC > kMaxOut)
    return set_error(ctx, QTNH_EUNSUPPORTED, "contract: %d closed / %d result axes exceed the direct kernel", P.nw,
                     P.rC);
  uint64_t stA[kMaxRank], stB[kMaxRank];
  strides(rA, dA, stA);
  strides(rB, dB, stB);
  Closed cl;
  memset(&cl, 0, sizeof(cl));
  cl.nw = P.nw;
  for (int w = 0; w < P.nw; ++w) {
    cl.kd[w] = dA[P.cA[w]];
    cl.ka[w] = stA[P.cA[w]];
    cl.kb[w] = stB[P.cB[w]];
  }
  if (P.K == 0) {
    // Empty sums: the result is zero.
    return launch_zero(ctx, C, P.nC * sizeof(c128));
  }
  // Natural axis q -> source stride (A or B); result axis p holds natural axis inv[p].
  const int n1 = (int)P.open1.size();
  std::vector<int> inv(P.rC);
  for (int q = 0; q < P.rC; ++q) inv[P.pos[q]] = q;
  BitLayout LA = bit_layout(rA, dA), LB = bit_layout(rB, dB), LC = bit_layout(P.rC, P.dC);
  if (LA.binary && LB.binary && LC.binary) {
    OutBits ob;
    memset(&ob, -1, sizeof(ob));
    ob.nbits = LC.nbits;
    for (int p = 0; p < P.rC; ++p) {
      int q = inv[p];
      for (int t = 0; t < LC.width[p]; ++t) {
        int cbit = LC.shift[p] + t;
        if (q < n1)
          ob.abit[cbit] = (int8_t)(LA.shift[P.open1[q]] + t);
        else
          ob.bbit[cbit] = (int8_t)(LB.shift[P.open2[q - n1]] + t);
      }
    }
    // Tuple bits, lowest first: the last closed position's bits, then the one before.
    ClosedBits cb;
    memset(&cb, 0, sizeof(cb));
    uint64_t fa = 0, fb = 0;
    for (int w = P.nw - 1; w >= 0; --w) {
      const int m = LA.width[P.cA[w]];
      for (int j = 0; j < m; ++j) {
        if (cb.kbits == kMaxClosedBits)
          return set_error(ctx, QTNH_EUNSUPPORTED, "contract: more than %d closed bits", kMaxClosedBits);
        fa |= 1ull << (LA.shift[P.cA[w]] + j);
        fb |= 1ull << (LB.shift[P.cB[w]] + j);
        cb.fa[cb.kbits] = fa;
        cb.fb[cb.kbits] = fb;
        ++cb.kbits;
      }
    }
    // Few B results on the lowest result bits and few terms (2^(NB + KB) <= 32):
    // one thread per row of B results.
    int nb_low = 0;
    while (nb_low < LC.nbits && ob.bbit[nb_low] >= 0) ++nb_low;
    bool b_low_only = nb_low >= 1;
    for (int p = nb_low; p < LC.nbits; ++p) b_low_only = b_low_only && ob.bbit[p] < 0;
    const uint64_t cap = (uint64_t)std::max(1, ctx->num_cus) * 16;
    // A square gate of 1 to 3 term bits (a dense gate contracted into a large
    // state with the default placement): the bit permutation that takes A's
    // term bits to the lowest result bits and its other bits to their result
    // bits, with each row of results formed in the permute's tiles.
    const uint64_t nA = 1ull << LA.nbits;
    if (b_low_only && cb.kbits == nb_low && nb_low <= 3 && LC.nbits >= 16 && LA.nbits == LC.nbits &&
        (C + P.nC <= A || A + nA <= C)) {
      int sigma[64];
      for (int p = nb_low; p < LC.nbits; ++p) sigma[ob.abit[p]] = p;
      for (int c = 0; c < cb.kbits; ++c) sigma[__builtin_ctzll(cb.fa[c] ^ (c ? cb.fa[c - 1] : 0ull))] = c;
      const int N = 1 << nb_low;
      uint32_t xb[64];
      for (int t = 0; t < N; ++t)
        for (int o = 0; o < N; ++o) {
          uint64_t x = 0;
          for (int c = 0; c < cb.kbits; ++c)
            if ((t >> c) & 1) x |= cb.fb[c] ^ (c ? cb.fb[c - 1] : 0ull);
          for (int j = 0; j < nb_low; ++j)
            if ((o >> j) & 1) x |= 1ull << ob.bbit[j];
          xb[t * N + o] = (uint32_t)x;
        }
      bool done = false;
      const qtnh_status st = launch_permute_gate(ctx, A, C, B, LC.nbits, sigma, nb_low, xb, &done);
      if (st || done) return st;
    }
    if (b_low_only && cb.kbits >= 1 && nb_low + cb.kbits <= kRowsMaxBits && LC.nbits - nb_low >= 12) {
      const uint64_t nrows = P.nC >> nb_low;
      const dim3 g((unsigned)cap);  // at most this many blocks, one per block step
      switch (nb_low) {
        case 1: launch_rows_kb<1, 4>(cb.kbits, g, ctx->stream, A, B, C, ob, cb, nrows); break;
        case 2: launch_rows_kb<2, 3>(cb.kbits, g, ctx->stream, A, B, C, ob, cb, nrows); break;
        case 3: launch_rows_kb<3, 2>(cb.kbits, g, ctx->stream, A, B, C, ob, cb, nrows); break;
        default: launch_rows_kb<4, 1>(cb.kbits, g, ctx->stream, A, B, C, ob, cb, nrows); break;
      }
      return check_hip(ctx, hipGetLastError(), "contract_direct_rows launch");
    }
    const int lowbits = std::min(LC.nbits, 8);
    const uint64_t nhi = P.nC >> lowbits;
    // Blocks loop over rows of 256 results, so their offset tables are built once per block.
    unsigned blocks = (unsigned)std::min<uint64_t>(nhi, cap);
    hipLaunchKernelGGL(contract_direct_bits, dim3(blocks), dim3(kDirThreads), 0, ctx->stream, A, B, C, ob, cb, P.nC);
    return check_hip(ctx, hipGetLastError(), "contract_direct_bits launch");
  }
  OutGen og;
  memset(&og, 0, sizeof(og));
  og.rc = P.rC;
  for (int p = 0; p < P.rC; ++p) {
    int q = inv[p];
    og.cd[p] = P.dC[p];
    if (q < n1)
      og.ca[p] = stA[P.open1[q]];
    else
      og.cb[p] = stB[P.open2[q - n1]];
  }
  unsigned blocks = grid_for(P.nC, kDirThreads, 1u << 20);
  hipLaunchKernelGGL(contract_direct_general, dim3(blocks), dim3(kDirThreads), 0, ctx->stream, A, B, C, og, cl, P.nC);
  return check_hip(ctx, hipGetLastError(), "contract_direct_general launch");
}

// Route choice: the direct kernel sums K operand pairs per result element in
// one thread, with no reuse, which is cheap while the sums are short (gates)
// or the whole contraction is small and its sums stay short.
static bool use_direct(const Plan &P) {
  if (P.nw > kMaxClosed || P.rC > kMaxOut) return false;
  return P.K <= 16 || (P.M * P.N * P.K <= (1ull << 20) && P.K <= 1024);
}

// Few outputs, long sums (inner products, closing steps of a network): the
// sums are split into ordered chunks over many workgroups.  Each output reads
// its own rows of A' and B', 2 M N K elements in all, where the GEMM route
// reads (M + N) K once but computes whole 64-wide tiles: per term, 32 M N
// bytes at ~5.5 TB/s against 8 x 64 x 64 flops at ~67 TFLOP/s plus 16 (M + N)
// bytes.  Split-K is the faster up to about 64 outputs (measured: 32 x 32
// outputs over 2^20 terms took 1.2 ms in split-K, reading 32 GiB of rows).
static bool use_splitk(const Plan &P) { return P.M * P.N <= 64; }

// Chunks per output of the split-K route; independent of the device.
static uint64_t splitk_chunks(uint64_t MN, uint64_t K) {
  const uint64_t want = (2048 + MN - 1) / MN;
  const uint64_t most = (K + 4095) / 4096;
  return std::max<uint64_t>(1, std::min(want, most));
}

// Operand and result layouts of the GEMM route and its workspace: A' = A
// permuted to [open1, closed]; B' = B permuted to [closed, open2] (FP64 GEMM)
// or [open2, closed] (split precision and split-K, which read B by columns);
// the natural result [open1, open2] is permuted to the placement.  Operands
// already in those layouts are used in place.  Workspace: A', B', natural C,
// then split scratch, split-K partials or GEMM depth-chunk partials.
static uint64_t pad_c128(uint64_t n) { return ((n * sizeof(c128)) + 255) & ~255ull; }

static void gemm_layout(const Plan &P, const uint64_t *dA, const uint64_t *dB, int precision, GemmRoute &L) {
  L.splitk = use_splitk(P);
  L.split = precision == QTNH_PREC_SPLIT && !L.splitk;
  L.M = P.M;
  L.N = P.N;
  L.K = P.K;
  L.nC = P.nC;
  L.pA.assign(P.rA, 0);
  L.pB.assign(P.rB, 0);
  L.pC.assign(P.rC, 0);
  int q = 0;
  for (int i : P.open1) L.pA[i] = (uint16_t)q++;
  for (int i : P.cA) L.pA[i] = (uint16_t)q++;
  // B: [open2, closed] for split-K; for split precision whichever of the
  // two layouts B already has ([open2, closed] or [closed, open2]: the split
  // reads either in place), else [closed, open2]; FP64 GEMM [closed, open2].
  auto layout_b = [&](bool open_first) {
    int qq = 0;
    if (open_first) {
      for (int i : P.open2) L.pB[i] = (uint16_t)qq++;
      for (int i : P.cB) L.pB[i] = (uint16_t)qq++;
    } else {
      for (int i : P.cB) L.pB[i] = (uint16_t)qq++;
      for (int i : P.open2) L.pB[i] = (uint16_t)qq++;
    }
    bool id = true;
    for (int i = 0; i < P.rB; ++i) id = id && L.pB[i] == i;
    return id;
  };
  L.b_cols = !L.splitk;
  if (L.splitk) {
    layout_b(true);
  } else if (L.split) {
    if (layout_b(true))
      L.b_cols = false;
    else
      layout_b(false);
  } else {
    layout_b(false);
  }
  L.idA = L.idB = L.idC = true;
  for (int i = 0; i < P.rA; ++i) L.idA = L.idA && L.pA[i] == i;
  for (int i = 0; i < P.rB; ++i) L.idB = L.idB && L.pB[i] == i;
  for (int i = 0; i < P.rC; ++i) {
    L.pC[i] = (uint16_t)P.pos[i];
    L.idC = L.idC && P.pos[i] == i;
  }
  L.dn.clear();
  for (int i : P.open1) L.dn.push_back(dA[i]);
  for (int i : P.open2) L.dn.push_back(dB[i]);
  L.kchunks = L.splitk ? splitk_chunks(P.M * P.N, P.K) : 0;
  L.zchunks = (L.split || L.splitk) ? 1 : zgemm_chunks(P.M, P.N, P.K);
  L.sbytes = L.split         ? ((split_scratch_bytes(P.M, P.N, P.K) + 255) & ~255ull)
             : L.splitk      ? pad_c128(P.M * P.N * L.kchunks)
             : L.zchunks > 1 ? pad_c128(P.M * P.N * L.zchunks)
                             : 0;
  uint64_t off = 0;
  L.offA = off;
  off += L.idA ? 0 : pad_c128(P.M * P.K);
  L.offB = off;
  off += L.idB ? 0 : pad_c128(P.K * P.N);
  L.offC = off;
  off += L.idC ? 0 : pad_c128(P.nC);
  L.offS = off;
  L.bytes = off + L.sbytes;
}

bool gemm_route(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                const uint16_t *r1, const uint16_t *r2, int precision, GemmRoute *out) {
  Plan P = make_plan(rA, dA, rB, dB, nw, wires, r1, r2);
  if (use_direct(P)) return false;
  gemm_layout(P, dA, dB, precision, *out);
  return true;
}

ContractRoute contract_route(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                             const uint16_t *r1, const uint16_t *r2, int precision) {
  Plan P = make_plan(rA, dA, rB, dB, nw, wires, r1, r2);
  ContractRoute R;
  R.M = P.M;
  R.N = P.N;
  R.K = P.K;
  R.direct = use_direct(P);
  R.ws_bytes = 0;
  if (!R.direct) {
    GemmRoute L;
    gemm_layout(P, dA, dB, precision, L);
    R.ws_bytes = L.bytes;
  }
  return R;
}

qtnh_status launch_contract(qtnh_ctx_t ctx, const c128 *A, int rA, const uint64_t *dA, const c128 *B, int rB,
                            const uint64_t *dB, int nw, const uint16_t *wires, const uint16_t *r1, const uint16_t *r2,
                            c128 *C, int precision, double *err_out, void *ext_ws, uint64_t ext_ws_bytes) {
  Plan P = make_plan(rA, dA, rB, dB, nw, wires, r1, r2);
  if (use_direct(P)) {
    // Small problems run exactly in FP64 whatever the precision asked for.
    if (err_out) *err_out = 0.0;
    return launch_contract_direct(ctx, A, rA, dA, B, rB, dB, nw, wires, r1, r2, C);
  }
  GemmRoute L;
  gemm_layout(P, dA, dB, precision, L);
  char *ws = nullptr;
  if (L.bytes && ext_ws) {
    // A caller-provided region (batches, captured graphs): never reallocated here.
    if (ext_ws_bytes < L.bytes)
      return set_error(ctx, QTNH_EINVAL, "contract: workspace of %llu bytes, need %llu",
                       (unsigned long long)ext_ws_bytes, (unsigned long long)L.bytes);
    ws = (char *)ext_ws;
  } else if (L.bytes) {
    void *w;
    qtnh_status st = workspace(ctx, L.bytes, &w);
    if (st) return st;
    ws = (char *)w;
  }
  const c128 *Ap = L.idA ? A : (const c128 *)(ws + L.offA);
  const c128 *Bp = L.idB ? B : (const c128 *)(ws + L.offB);
  c128 *Cn = L.idC ? C : (c128 *)(ws + L.offC);
  if (!L.idA) {
    qtnh_status st = launch_permute(ctx, A, (c128 *)Ap, rA, dA, L.pA.data());
    if (st) return st;
  }
  if (!L.idB) {
    qtnh_status st = launch_permute(ctx, B, (c128 *)Bp, rB, dB, L.pB.data());
    if (st) return st;
  }
  qtnh_status st;
  if (L.splitk) {
    st = launch_zgemm_splitk(ctx, Ap, Bp, Cn, P.M, P.N, P.K, L.kchunks, (c128 *)(ws + L.offS));
    if (err_out) *err_out = 0.0;
  } else if (L.split) {
    st = launch_zgemm_split(ctx, Ap, Bp, L.b_cols, Cn, P.M, P.N, P.K, ws + L.offS, L.sbytes, err_out);
  } else {
    st = launch_zgemm_f64(ctx, Ap, Bp, Cn, P.M, P.N, P.K, P.K, P.N, P.N,
                          L.zchunks > 1 ? (c128 *)(ws + L.offS) : nullptr);
    if (err_out) *err_out = 0.0;
  }
  if (st) return st;
  if (!L.idC) {
    st = launch_permute(ctx, Cn, C, P.rC, L.dn.data(), L.pC.data());
    if (st) return st;
  }
  return QTNH_OK;
}

}  // namespace qtnh_hip
